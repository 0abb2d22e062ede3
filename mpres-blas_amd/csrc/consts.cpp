// Host-side computation of the per-moduli-set constants (DevConsts).
//
// Follows rns_const_init (src/rns.cuh:324-442) and mp_const_init
// (src/arith/arith_utils.cuh:44-85) without GMP/MPFR:
//   * M = prod m_i is formed in 32-bit limbs; every floating-point use sees
//     M rounded down to 53 bits (the reference's MPFR variable has the
//     default 53-bit precision and is set with MPFR_RNDD, src/rns.cuh:326,336);
//   * RNS_EVAL_UNIT / INV_UNIT are directed roundings of 1/M53 and 1-1/M53 with
//     MPFR's frac-in-[0.5,1) normalisation (src/rns.cuh:388-411);
//   * the directed reciprocals use the emulated rounding of
//     src/dinterval.cuh:142-154 (src/rns.cuh:412-415), as the GPU path of the
//     reference copies these host values to the device.
#include <cfenv>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "devconsts.hpp"
#include "internal.hpp"

namespace mph {

// The moduli sets of the reference params files with n in {8, 16, 32}:
// src/params/<dir>/params.<name>.h (RNS_MODULI_VALUES, RNS_MODULI_PRODUCT_LOG2).
static const ModuliSet SETS[] = {
    {"8_2double", 8, 214, {113812103,113812105,113812107,113812109,113812111,113812117,113812121,113812123}},
    {"16_4double", 16, 426, {103812103,103812105,103812107,103812109,103812113,103812119,103812127,103812131,
                             103812133,103812137,103812139,103812151,103812157,103812161,103812167,103812169}},
    {"32_8double", 32, 850, {99312103,99312105,99312107,99312109,99312113,99312119,99312121,99312127,
                             99312131,99312133,99312137,99312139,99312143,99312149,99312151,99312163,
                             99312167,99312173,99312179,99312193,99312197,99312203,99312211,99312217,
                             99312223,99312233,99312247,99312253,99312263,99312271,99312277,99312287}},
    {"8_32bit", 8, 242, {1283742825,1283742827,1283742829,1283742833,1283742839,1283742841,1283742847,1283742851}},
    {"16_32bit", 16, 482, {1173342131,1173342133,1173342135,1173342137,1173342139,1173342143,1173342151,1173342161,
                           1173342167,1173342173,1173342179,1173342187,1173342193,1173342197,1173342199,1173342211}},
    {"32_32bit", 32, 962, {1122512117,1122512119,1122512121,1122512123,1122512125,1122512129,1122512131,1122512137,
                           1122512141,1122512143,1122512147,1122512149,1122512161,1122512173,1122512177,1122512179,
                           1122512189,1122512191,1122512197,1122512201,1122512203,1122512207,1122512213,1122512219,
                           1122512221,1122512231,1122512233,1122512239,1122512243,1122512249,1122512257,1122512269}},
    {"8_16bit", 8, 122, {39051,39053,39055,39059,39061,39071,39073,39077}},
    {"16_16bit", 16, 242, {36051,36053,36055,36059,36061,36067,36071,36073,36077,36079,36083,36089,36091,36097,36101,36103}},
    {"32_16bit", 32, 482, {34403,34405,34407,34409,34411,34417,34421,34423,34427,34429,34439,34441,34451,34453,34457,34459,
                           34469,34471,34477,34483,34487,34499,34501,34511,34513,34519,34537,34543,34547,34549,34553,34571}},
};

int num_sets() { return (int)(sizeof(SETS) / sizeof(SETS[0])); }
const ModuliSet *get_set(int id) { return (id >= 0 && id < num_sets()) ? &SETS[id] : nullptr; }

static int64_t inverse_mod(int64_t a, int64_t m) {
    int64_t t = 0, nt = 1, r = m, nr = ((a % m) + m) % m;
    while (nr != 0) {
        int64_t q = r / nr, tmp;
        tmp = t - q * nt; t = nt; nt = tmp;
        tmp = r - q * nr; r = nr; nr = tmp;
    }
    return t < 0 ? t + m : t;
}

// emulated directed division of src/dinterval.cuh:142-154 (fmac unfused)
static double emu_div_rd(double a, double b) {
    volatile double c = a / b;
    volatile double e = 0x1.0000000000001p-53 * std::fabs(c);
    e = e + 0x1p-1074;
    return c - e;
}
static double emu_div_ru(double a, double b) {
    volatile double c = a / b;
    volatile double e = 0x1.0000000000001p-53 * std::fabs(c);
    e = e + 0x1p-1074;
    return c + e;
}

bool compute_consts(const ModuliSet &s, DevConsts &c, std::vector<int> &pow2) {
    std::memset(&c, 0, sizeof c);
    const int n = s.n;
    c.n = n;
    c.log2m = s.log2m;
    for (int i = 0; i < n; i++) {
        c.m[i] = s.m[i];
        c.md[i] = (double)s.m[i];
        c.minv[i] = 1.0 / (double)s.m[i];
    }
    // M in 32-bit limbs
    std::vector<uint32_t> M(1, 1u);
    for (int i = 0; i < n; i++) {
        uint64_t carry = 0;
        for (auto &limb : M) {
            uint64_t t = (uint64_t)limb * (uint32_t)s.m[i] + carry;
            limb = (uint32_t)t;
            carry = t >> 32;
        }
        if (carry) M.push_back((uint32_t)carry);
    }
    const int bitlen = 32 * ((int)M.size() - 1) + (32 - __builtin_clz(M.back()));
    if (bitlen - 1 != s.log2m) return false;  // the params files' RNS_MODULI_PRODUCT_LOG2 is floor(log2 M)
    const int sh = bitlen - 53;
    uint64_t K = 0;
    for (int b = 52; b >= 0; b--) {
        const int pos = sh + b;
        K = (K << 1) | ((M[pos / 32] >> (pos % 32)) & 1u);
    }
    for (int i = 0; i < n; i++) {
        int64_t p = 1;
        for (int j = 0; j < n; j++)
            if (j != i) p = (p * (s.m[j] % s.m[i])) % s.m[i];
        c.w[i] = (int)inverse_mod(p, s.m[i]);
    }
    pow2.assign((size_t)(s.log2m + 1) * n, 0);
    for (int j = 0; j < n; j++) {
        int64_t v = 1;
        for (int i = 0; i <= s.log2m; i++) {
            pow2[(size_t)i * n + j] = (int)v;
            v = (v * 2) % s.m[j];
        }
    }
    uint32_t Mlow = 1;
    for (int j = 0; j < n; j++) Mlow *= (uint32_t)s.m[j];
    for (int i = 0; i < P2T; i++) {
        const uint32_t mask = (1u << (i + 1)) - 1u;
        c.mpow2[i] = (int)(Mlow & mask);
        for (int j = 0; j < n; j++) {
            uint32_t p = 1;
            for (int k = 0; k < n; k++)
                if (k != j) p *= (uint32_t)s.m[k];
            c.h[i][j] = (int)(p & mask);
            int64_t q = 1;
            for (int k = 0; k <= i; k++) q = (q * 2) % s.m[j];
            c.pow2inv[i][j] = (int)inverse_mod(q, s.m[j]);
        }
    }
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) c.mrcinv[i][j] = (j > i) ? (int)inverse_mod(s.m[i], s.m[j]) : 0;
    // accuracy constant and refinement factor, src/rns.cuh:385-387 (same order of double operations)
    volatile double acc = 4 * std::pow(2.0, 1 - 53);
    acc = acc * n;
    acc = acc * std::log2((double)n);
    acc = acc * (1 + (0.0000001) / 2);
    acc = acc / (0.0000001);
    c.acc = acc;
    c.ref_factor = (int)std::floor(std::log2(1 / (2 * c.acc)));
    // 1/M53 = 2^-sh / K, directed-rounded, as (frac in [0.5,1), exp)
    {
        const int old = std::fegetround();
        volatile double Kd = (double)K;
        std::fesetround(FE_DOWNWARD);
        volatile double lo = 1.0 / Kd;
        std::fesetround(FE_UPWARD);
        volatile double hi = 1.0 / Kd;
        std::fesetround(old);
        int e;
        double f = std::frexp(lo, &e);
        c.unit_low = {f, e - sh, 0};
        f = std::frexp(hi, &e);
        c.unit_upp = {f, e - sh, 0};
        f = std::frexp(1.0, &e);
        c.inv_unit_upp = {f, e, 0};
        // two-double reciprocals of the unit's fractions (rounding to nearest)
        c.ul_rh = 1.0 / c.unit_low.frac;
        c.ul_rl = std::fma(-c.ul_rh, c.unit_low.frac, 1.0) / c.unit_low.frac;
        c.uu_rh = 1.0 / c.unit_upp.frac;
        c.uu_rl = std::fma(-c.uu_rh, c.unit_upp.frac, 1.0) / c.unit_upp.frac;
        f = std::frexp(1.0 - std::ldexp(1.0, -53), &e);
        c.inv_unit_low = {f, e, 0};
    }
    for (int i = 0; i < n; i++) {
        c.rrd[i] = emu_div_rd(1.0, (double)s.m[i]);
        c.rru[i] = emu_div_ru(1.0, (double)s.m[i]);
    }
    // MP_PRECISION, MP_H, MP_J from b = floor(log2 M53) = floor(log2 M)
    const int b = bitlen - 1;
    c.mp_precision = b / 2 - 1;
    c.mp_h = -(b / 2) - 1;
    c.mp_j = -1;
    c.pow2 = nullptr;
    return true;
}

}  // namespace mph
