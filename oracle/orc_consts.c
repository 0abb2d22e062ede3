/*
 * Oracle: moduli sets and the precomputed RNS / mp constants.
 *
 * TEST INFRASTRUCTURE ONLY (see mpres_oracle.h).
 *
 * Restates rns_const_init (src/rns.cuh:324-442) and mp_const_init
 * (src/arith/arith_utils.cuh:44-85) without GMP/MPFR: the only multi-word
 * quantity needed is M = prod m_i, kept here as little-endian 32-bit limbs.
 *   - M is rounded down to 53 bits before any floating-point use, because the
 *     reference stores it in an MPFR variable of default (53-bit) precision
 *     with MPFR_RNDD (src/rns.cuh:326,336).
 *   - RNS_EVAL_UNIT / RNS_EVAL_INV_UNIT keep MPFR's mpfr_get_d_2exp
 *     normalisation, frac in [0.5, 1) (src/rns.cuh:396-407).
 *   - The moduli reciprocals are rounded with the emulated (policy E) ddiv_rd
 *     and ddiv_ru of src/dinterval.cuh:142-154 (src/rns.cuh:412-415); the GPU
 *     path uses these same values.
 */
#include <fenv.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "orc_internal.h"

/* Moduli sets: the values of the reference params files with n in {8,16,32}
 * (src/params/<dir>/params.<name>.h, RNS_MODULI_VALUES and
 * RNS_MODULI_PRODUCT_LOG2). */
static const struct { const char *name; int n; int log2m; int m[32]; } SETS[] = {
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
#define NSETS ((int)(sizeof(SETS) / sizeof(SETS[0])))

int orc_num_sets(void) { return NSETS; }
const char *orc_set_name(int id) { return (id >= 0 && id < NSETS) ? SETS[id].name : 0; }

/* Modular inverse by the extended Euclidean algorithm (textbook; the
 * reference's mod_inverse, src/rns.cuh:140-172, computes the same value). */
static int inv_mod(long a, long m) {
    long t = 0, nt = 1, r = m, nr = a % m;
    if (nr < 0) nr += m;
    while (nr != 0) {
        long q = r / nr, tmp;
        tmp = t - q * nt; t = nt; nt = tmp;
        tmp = r - q * nr; r = nr; nr = tmp;
    }
    if (t < 0) t += m;
    return (int)t;
}

/* Policy-E directed division, src/dinterval.cuh:142-154 (fmac unfused). */
static double e_ddiv_rd(double a, double b) {
    double c = a / b;
    return c - (ORC_PHI1 * fabs(c) + ORC_ETA);
}
static double e_ddiv_ru(double a, double b) {
    double c = a / b;
    return c + (ORC_PHI1 * fabs(c) + ORC_ETA);
}

orc_ctx *orc_ctx_create(int id) {
    if (id < 0 || id >= NSETS) return 0;
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    const int n = SETS[id].n;
    c->n = n;
    c->log2m = SETS[id].log2m;
    for (int i = 0; i < n; i++) c->m[i] = SETS[id].m[i];

    /* M as a multi-word integer */
    uint32_t limb[40];
    int nl = 1;
    memset(limb, 0, sizeof limb);
    limb[0] = 1;
    for (int i = 0; i < n; i++) {
        uint64_t carry = 0;
        for (int k = 0; k < nl; k++) {
            uint64_t t = (uint64_t)limb[k] * (uint32_t)c->m[i] + carry;
            limb[k] = (uint32_t)t;
            carry = t >> 32;
        }
        if (carry) limb[nl++] = (uint32_t)carry;
    }
    int top = 31;
    while (!((limb[nl - 1] >> top) & 1)) top--;
    const int bitlen = 32 * (nl - 1) + top + 1;
    c->log2m_floor = bitlen - 1;
    /* K = top 53 bits of M (M rounded down to 53 bits = K * 2^s) */
    const int s = bitlen - 53;
    uint64_t K = 0;
    for (int b = 52; b >= 0; b--) {
        int pos = s + b;
        K = (K << 1) | ((limb[pos / 32] >> (pos % 32)) & 1);
    }

    /* w_i = (M / m_i)^-1 mod m_i  (src/rns.cuh:343-346) */
    for (int i = 0; i < n; i++) {
        long p = 1;
        for (int j = 0; j < n; j++)
            if (j != i) p = (p * (c->m[j] % c->m[i])) % c->m[i];
        c->w[i] = inv_mod(p, c->m[i]);
    }
    /* RNS_POW2[i][j] = 2^i mod m_j, i = 0..L (src/rns.cuh:352-358) */
    c->pow2 = (int *)malloc(sizeof(int) * (size_t)(c->log2m + 1) * n);
    for (int j = 0; j < n; j++) {
        long v = 1;
        for (int i = 0; i <= c->log2m; i++) {
            c->pow2[i * n + j] = (int)v;
            v = (v * 2) % c->m[j];
        }
    }
    /* M mod 2^(i+1) and M_j mod 2^(i+1) (src/rns.cuh:362-372): products mod 2^32 */
    uint32_t Mw = 1;
    for (int j = 0; j < n; j++) Mw *= (uint32_t)c->m[j];
    for (int i = 0; i < 30; i++) {
        uint32_t mask = (i + 1 == 32) ? 0xffffffffu : ((1u << (i + 1)) - 1u);
        c->mpow2[i] = (int)(Mw & mask);
        for (int j = 0; j < n; j++) {
            uint32_t p = 1;
            for (int k = 0; k < n; k++)
                if (k != j) p *= (uint32_t)c->m[k];
            c->h[i][j] = (int)(p & mask);
        }
    }
    /* (2^(i+1))^-1 mod m_j (src/rns.cuh:374-382) */
    for (int i = 0; i < 30; i++)
        for (int j = 0; j < n; j++) {
            long p = 1;
            for (int k = 0; k <= i; k++) p = (p * 2) % c->m[j];
            c->pow2inv[i][j] = inv_mod(p, c->m[j]);
        }
    /* Accuracy constant and refinement factor (src/rns.cuh:385-387), evaluated
     * in the same order of double operations. */
    c->eval_acc = 4 * pow(2.0, 1 - 53) * n * log2((double)n) * (1 + (0.0000001) / 2) / (0.0000001);
    c->ref_factor = (int)floor(log2(1 / (2 * c->eval_acc)));
    /* RNS_EVAL_UNIT = [RD(1/M53), RU(1/M53)], RNS_EVAL_INV_UNIT = [RD(1-RD(1/M53)), RU(1-RU(1/M53))]
     * as (frac in [0.5,1), exp) pairs (src/rns.cuh:388-411). 1/M53 = 2^-s / K. */
    {
        volatile double Kd = (double)K;
        int e;
        fesetround(FE_DOWNWARD);
        volatile double lo = 1.0 / Kd;
        fesetround(FE_UPWARD);
        volatile double hi = 1.0 / Kd;
        fesetround(FE_TONEAREST);
        double f = frexp(lo, &e);
        c->unit_low.frac = f;
        c->unit_low.exp = e - s;
        f = frexp(hi, &e);
        c->unit_upp.frac = f;
        c->unit_upp.exp = e - s;
        /* 1 - RU(1/M53) rounded up to 53 bits is 1; 1 - RD(1/M53) rounded down is 1 - 2^-53 */
        f = frexp(1.0, &e);
        c->inv_unit_upp.frac = f;
        c->inv_unit_upp.exp = e;
        f = frexp(1.0 - ldexp(1.0, -53), &e);
        c->inv_unit_low.frac = f;
        c->inv_unit_low.exp = e;
    }
    /* directed reciprocals, policy E (src/rns.cuh:412-415) */
    for (int i = 0; i < n; i++) {
        c->recip_rd[i] = e_ddiv_rd(1.0, (double)c->m[i]);
        c->recip_ru[i] = e_ddiv_ru(1.0, (double)c->m[i]);
    }
    /* MRC_MULT_INV[i][j] = m_i^-1 mod m_j for j > i, else 0 (src/rns.cuh:417-426) */
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++)
            c->mrc_inv[i][j] = (j > i) ? inv_mod(c->m[i], c->m[j]) : 0;
    /* MP_PRECISION, MP_H, MP_J (src/arith/arith_utils.cuh:44-61) from
     * b = floor(log2 M53) = floor(log2 M): floor(log2 sqrt(M53)) = floor(b/2),
     * floor(log2(sqrt(M53)/M53)) = -floor(b/2) - 1, floor(log2(1 - M53^-1/2)) = -1.
     * (M53 is never a power of two for these sets.) */
    c->mp_precision = c->log2m_floor / 2 - 1;
    c->mp_h = -(c->log2m_floor / 2) - 1;
    c->mp_j = -1;
    return c;
}

void orc_ctx_destroy(orc_ctx *c) {
    if (!c) return;
    free(c->pow2);
    free(c);
}

int orc_n(const orc_ctx *c) { return c->n; }
int orc_log2m(const orc_ctx *c) { return c->log2m; }
void orc_moduli(const orc_ctx *c, int *o) { memcpy(o, c->m, sizeof(int) * c->n); }
void orc_part_inverse(const orc_ctx *c, int *o) { memcpy(o, c->w, sizeof(int) * c->n); }
void orc_pow2_table(const orc_ctx *c, int *o) { memcpy(o, c->pow2, sizeof(int) * (size_t)(c->log2m + 1) * c->n); }
void orc_pow2_residues(const orc_ctx *c, int *o) { memcpy(o, c->mpow2, sizeof(int) * 30); }
void orc_part_pow2_residues(const orc_ctx *c, int *o) {
    for (int i = 0; i < 30; i++) memcpy(o + i * c->n, c->h[i], sizeof(int) * c->n);
}
void orc_pow2_inverse(const orc_ctx *c, int *o) {
    for (int i = 0; i < 30; i++) memcpy(o + i * c->n, c->pow2inv[i], sizeof(int) * c->n);
}
void orc_mrc_inverse(const orc_ctx *c, int *o) {
    for (int i = 0; i < c->n; i++) memcpy(o + i * c->n, c->mrc_inv[i], sizeof(int) * c->n);
}
void orc_recip(const orc_ctx *c, double *rd, double *ru) {
    memcpy(rd, c->recip_rd, sizeof(double) * c->n);
    memcpy(ru, c->recip_ru, sizeof(double) * c->n);
}
void orc_eval_consts(const orc_ctx *c, double *acc, int *ref, orc_er_t *ul, orc_er_t *uu, orc_er_t *il, orc_er_t *iu) {
    *acc = c->eval_acc;
    *ref = c->ref_factor;
    *ul = c->unit_low;
    *uu = c->unit_upp;
    *il = c->inv_unit_low;
    *iu = c->inv_unit_upp;
}
void orc_mp_consts(const orc_ctx *c, int *p, int *h, int *j) {
    *p = c->mp_precision;
    *h = c->mp_h;
    *j = c->mp_j;
}
