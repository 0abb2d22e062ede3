// Scalar numerics shared by every kernel of the RNS multiple-precision path:
// IEEE-754 directed rounding, extended-range floats, exact modular arithmetic.
//
// Everything here is __host__ __device__ so the same code is checked on the
// CPU (tests/test_numerics_host.py builds a host-only harness of this header)
// before it runs on the GPU.
//
// Directed rounding. The reference's CUDA path uses __dadd_rd / __dmul_ru /
// __ddiv_rd ... (true IEEE directed rounding, e.g. src/extrange.cuh:458-641,
// src/rns.cuh:807-862). gfx950 has no per-instruction rounding mode; on the GPU
// a directed add, multiply or fma runs with the MODE register's FP64 round field
// set around that one instruction (see hw_add below). On the host, and as the
// definition those instructions are checked against, each directed operation is
// derived from the round-to-nearest result and its exact error term:
//   a + b = s + e   (Knuth's TwoSum, exact for finite inputs)
//   a * b = p + e   (e = fma(a, b, -p), exact without underflow)
//   a / b = q + r/b (r = fma(-q, b, a), exact without underflow)
// and the rounded-to-nearest value is stepped one ulp down (up) when the
// error shows it lies above (below) the exact result. Exact zero sums follow
// IEEE-754: -0 under round-down unless both operands are +0.
//
// Modular arithmetic. Residues are < 2^31 and products < 2^62. Products are
// split into a round-to-nearest double p and its exact low part; the quotient
// floor(p / m) is estimated from p * (1/m) and the remainder p - q*m is
// formed exactly by an fma, then corrected into [0, m). The result is the
// canonical residue, i.e. what the reference's 64-bit '%' gives
// (src/modular.cuh:118-174, 250-314). All of this is full-rate FP64 on CDNA4.
#pragma once

#include <stdint.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MPH_HD __host__ __device__ __forceinline__
#else
#define MPH_HD inline __attribute__((always_inline))
#endif

namespace mph {

MPH_HD uint64_t d2u(double x) {
    union { double d; uint64_t u; } v;
    v.d = x;
    return v.u;
}
MPH_HD double u2d(uint64_t x) {
    union { double d; uint64_t u; } v;
    v.u = x;
    return v.d;
}

// next double toward -inf / +inf of a finite nonzero x
MPH_HD double step_down(double x) { return u2d(x > 0 ? d2u(x) - 1 : d2u(x) + 1); }
MPH_HD double step_up(double x) { return u2d(x > 0 ? d2u(x) + 1 : d2u(x) - 1); }

#if defined(__HIP_DEVICE_COMPILE__)
// On the GPU the FP64 round-mode field of the MODE register (bits 3:2: 0 to nearest even,
// 1 toward +inf, 2 toward -inf) gives IEEE directed rounding directly. Each block writes
// the field, runs its one instruction and writes round-to-nearest back, with two wait
// states after every write, so no instruction outside the block runs under a directed
// mode and the block can be scheduled like any other instruction.
#define MPH_DIRECTED(MODE, INSN)                                                              \
    "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), " #MODE "\n\ts_nop 1\n\t" INSN "\n\ts_nop 1\n\t" \
    "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 0\n\ts_nop 1"
__device__ __forceinline__ double hw_add(double a, double b, bool up) {
    double r;
    if (up)
        asm(MPH_DIRECTED(1, "v_add_f64 %0, %1, %2") : "=v"(r) : "v"(a), "v"(b));
    else
        asm(MPH_DIRECTED(2, "v_add_f64 %0, %1, %2") : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ double hw_mul(double a, double b, bool up) {
    double r;
    if (up)
        asm(MPH_DIRECTED(1, "v_mul_f64 %0, %1, %2") : "=v"(r) : "v"(a), "v"(b));
    else
        asm(MPH_DIRECTED(2, "v_mul_f64 %0, %1, %2") : "=v"(r) : "v"(a), "v"(b));
    return r;
}
__device__ __forceinline__ double hw_fma(double a, double b, double c, bool up) {
    double r;
    if (up)
        asm(MPH_DIRECTED(1, "v_fma_f64 %0, %1, %2, %3") : "=v"(r) : "v"(a), "v"(b), "v"(c));
    else
        asm(MPH_DIRECTED(2, "v_fma_f64 %0, %1, %2, %3") : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}
// Pairs: the lower bound's operation rounded down and the upper bound's rounded up in
// one block (three mode writes for two instructions). The lower result is written
// before the upper operands are read, hence the early-clobber output.
#define MPH_PAIR(INSN_LO, INSN_HI)                                                          \
    "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 2\n\ts_nop 1\n\t" INSN_LO "\n\ts_nop 0\n\t"     \
    "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 1\n\ts_nop 1\n\t" INSN_HI "\n\ts_nop 0\n\t"     \
    "s_setreg_imm32_b32 hwreg(HW_REG_MODE, 2, 2), 0\n\ts_nop 1"
__device__ __forceinline__ void hw_add_pair(double a, double b, double c, double d, double &lo, double &hi) {
    asm(MPH_PAIR("v_add_f64 %0, %2, %3", "v_add_f64 %1, %4, %5")
        : "=&v"(lo), "=v"(hi)
        : "v"(a), "v"(b), "v"(c), "v"(d));
}
__device__ __forceinline__ void hw_mul_pair(double a, double b, double c, double d, double &lo, double &hi) {
    asm(MPH_PAIR("v_mul_f64 %0, %2, %3", "v_mul_f64 %1, %4, %5")
        : "=&v"(lo), "=v"(hi)
        : "v"(a), "v"(b), "v"(c), "v"(d));
}
__device__ __forceinline__ void hw_fma_pair(double a, double b, double c, double d, double e, double f, double &lo,
                                            double &hi) {
    asm(MPH_PAIR("v_fma_f64 %0, %2, %3, %4", "v_fma_f64 %1, %5, %6, %7")
        : "=&v"(lo), "=v"(hi)
        : "v"(a), "v"(b), "v"(c), "v"(d), "v"(e), "v"(f));
}
#undef MPH_PAIR
#undef MPH_DIRECTED
#endif

// On the host (and in the host tests of this header) the same results come from the
// error-free transforms described at the top of this file.
MPH_HD double add_rd(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return hw_add(a, b, false);
#else
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    if (e < 0) return step_down(s);
    if (s == 0) return (d2u(a) == 0 && d2u(b) == 0) ? 0.0 : -0.0;
    return s;
#endif
}
MPH_HD double add_ru(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return hw_add(a, b, true);
#else
    double s = a + b;
    double bb = s - a;
    double e = (a - (s - bb)) + (b - bb);
    return (e > 0) ? step_up(s) : s;
#endif
}
MPH_HD double sub_rd(double a, double b) { return add_rd(a, -b); }
MPH_HD double sub_ru(double a, double b) { return add_ru(a, -b); }
MPH_HD double mul_rd(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return hw_mul(a, b, false);
#else
    double p = a * b;
    double e = fma(a, b, -p);
    return (e < 0) ? step_down(p) : p;
#endif
}
MPH_HD double mul_ru(double a, double b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return hw_mul(a, b, true);
#else
    double p = a * b;
    double e = fma(a, b, -p);
    return (e > 0) ? step_up(p) : p;
#endif
}
// lo = a * b rounded down, hi = c * d rounded up
MPH_HD void mul_rd_ru(double a, double b, double c, double d, double &lo, double &hi) {
#if defined(__HIP_DEVICE_COMPILE__)
    hw_mul_pair(a, b, c, d, lo, hi);
#else
    lo = mul_rd(a, b);
    hi = mul_ru(c, d);
#endif
}
MPH_HD double div_rd(double a, double b) {
    double q = a / b;
    double r = fma(-q, b, a);
    bool below = (r != 0) && ((r < 0) != (b < 0));  // exact a/b < q
    return below ? step_down(q) : q;
}
MPH_HD double div_ru(double a, double b) {
    double q = a / b;
    double r = fma(-q, b, a);
    bool above = (r != 0) && ((r < 0) == (b < 0));  // exact a/b > q
    return above ? step_up(q) : q;
}

// ceil(log2(x)) where log2 is rounded to nearest, for a positive normal x.
// log2(x) = e + log2(f), f in [1,2). It rounds to e exactly when f = 1 or
// log2(f) is below half the spacing of doubles just above e (toward e+1).
MPH_HD int ceil_log2_rn(double x) {
    const uint64_t u = d2u(x);
    const int e = (int)((u >> 52) & 2047) - 1023;
    const uint64_t mant = u & ((((uint64_t)1) << 52) - 1);
    if (mant == 0) return e;
    if (e == 0) return 1;
    const int a = e > 0 ? e : -e;
    int fl = 31 - __builtin_clz((unsigned)a);         // floor(log2 |e|)
    if (e < 0 && (a & (a - 1)) == 0) fl -= 1;         // just above -2^k lies the binade below
    // half spacing = 2^(fl-53); log2(f) < 2^(fl-53)  <=>  mant < 2^(fl-1) * ln 2  (to first order)
    const double t = ldexp(0.69314718055994530942, fl - 1);
    return ((double)mant < t) ? e : e + 1;
}

// ---------------------------------------------------------------------------
// Extended-range floats (er_float_t: frac in [1,2) and an integer exponent),
// src/extrange.cuh. Exponents are held in int registers.
// ---------------------------------------------------------------------------
struct Er {
    double frac;
    int exp;
};

// er_adjust, src/extrange.cuh:412-421: frac scaled into [1, 2) (sign kept), exp
// corrected; a zero frac gives exp 0. The reference reads the exponent field; here
// frexp does it (v_frexp_mant / v_frexp_exp on the GPU), which is the same for every
// input this code passes: normal doubles (products, quotients and sums of fractions in
// [1/2, 2), or of a fraction and a 53-bit significand) and signed zeros.
MPH_HD Er er_adjust(double frac, int exp) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double m = __builtin_amdgcn_frexp_mant(frac);
    const int e = __builtin_amdgcn_frexp_exp(frac);
#else
    int e;
    const double m = frexp(frac, &e);
#endif
    Er r;
    r.frac = m + m;
    r.exp = (frac != 0) ? exp + e - 1 : 0;
    return r;
}
// er_set_d, src/extrange.cuh:46-58 (CUDA :94-106)
MPH_HD Er er_set_d(double x) {
    Er r;
    if (x != 0) {
        const uint64_t u = d2u(x);
        r.exp = (int)((u & ~((uint64_t)1 << 63)) >> 52) - 1023;
        r.frac = u2d((u & ((((uint64_t)1) << 52) - 1)) | (u & ((uint64_t)1 << 63)) | ((uint64_t)1023 << 52));
    } else {
        r.exp = 0;
        r.frac = 0;
    }
    return r;
}
// er_ucmp (CUDA), src/extrange.cuh:671-677: comparison of the magnitudes of two
// extended-range numbers with nonnegative fractions
MPH_HD int er_ucmp(Er x, Er y) {
    const int diff = x.exp - y.exp;
    if (diff == 0) return x.frac == y.frac ? 0 : (x.frac > y.frac ? 1 : -1);
    return (x.frac != 0 && (diff > 0 || (diff < 0 && y.frac == 0))) ? 1 : -1;
}
// er_add_rd / er_add_ru (CUDA), src/extrange.cuh:450-490: the operands aligned to the
// larger exponent e, then one directed addition and er_adjust
// (written without branches: for dexp = 0, ldexp(y.frac, 0) = y.frac and e is x.exp
// unless x.exp is 0)
MPH_HD void er_add_align(Er x, Er y, double &a, double &b, int &e) {
    const int xe = (x.frac != 0) ? x.exp : 0, ye = (y.frac != 0) ? y.exp : 0;
    const int dexp = (x.frac != 0 && y.frac != 0) ? xe - ye : 0;
    const bool xg = dexp >= 0;
    a = xg ? x.frac : y.frac;
    b = ldexp(xg ? y.frac : x.frac, xg ? -dexp : dexp);
    e = xg ? ((dexp == 0 && xe == 0) ? ye : xe) : ye;
}
MPH_HD Er er_add_dir(Er x, Er y, bool up) {
    double a, b;
    int e;
    er_add_align(x, y, a, b, e);
    const double f = up ? add_ru(a, b) : add_rd(a, b);
    return er_adjust(f, e);
}
// lo = er_add_rd(xl, yl), hi = er_add_ru(xu, yu)
MPH_HD void er_add_pair(Er xl, Er yl, Er xu, Er yu, Er &lo, Er &hi) {
    double al, bl, au, bu;
    int el, eu;
    er_add_align(xl, yl, al, bl, el);
    er_add_align(xu, yu, au, bu, eu);
#if defined(__HIP_DEVICE_COMPILE__)
    double fl, fu;
    hw_add_pair(al, bl, au, bu, fl, fu);
#else
    const double fl = add_rd(al, bl), fu = add_ru(au, bu);
#endif
    lo = er_adjust(fl, el);
    hi = er_adjust(fu, eu);
}
// er_md_rd / er_md_ru: x*y/z, src/extrange.cuh:627-644
MPH_HD Er er_md_rd(Er x, Er y, Er z) { return er_adjust(div_rd(mul_rd(x.frac, y.frac), z.frac), x.exp + y.exp - z.exp); }
MPH_HD Er er_md_ru(Er x, Er y, Er z) { return er_adjust(div_ru(mul_ru(x.frac, y.frac), z.frac), x.exp + y.exp - z.exp); }

// Directed division by a fixed z in [2^-4, 2^4], for 2^-900 < |a| < 2^900 or a = 0, from a
// two-double reciprocal zh + zl of z (recip2): zh = RN(1/z) and zl = RN(fma(-zh, z, 1) / z)
// (the fma is exact), so |1/z - zh - zl| <= 2^-105 / z. q = RN(a zh + RN(a zl)) rounds a value
// within 2^-104 |a/z| of a/z: no double lies strictly between them unless a/z is within that
// distance of it, so q is one of the two doubles enclosing a/z (or a/z itself). The remainder
// r = a - q z is exact (fma; q is within one ulp of a/z) and its sign tells on which side q
// lies. The results equal div_rd(a, z) / div_ru(a, z) without the division sequence.
MPH_HD void recip2(double z, double &zh, double &zl) {
    zh = 1.0 / z;
    zl = fma(-zh, z, 1.0) / z;
}
// On the GPU the step is one directed fma: q + r 2^-64 lies strictly between q and its
// neighbour on r's side when r != 0 (|r 2^-64| is far below half an ulp of q, and the fma
// forms the sum exactly before its one rounding), so rounding it down (up) gives q's lower
// (upper) neighbour exactly when q lies above (below) a/z, and q itself otherwise.
MPH_HD double divk_rd(double a, double z, double zh, double zl) {
    const double q = fma(a, zh, a * zl);
    const double r = fma(-q, z, a);
#if defined(__HIP_DEVICE_COMPILE__)
    return (a == 0) ? a : hw_fma(r, 0x1p-64, q, false);  // a zero quotient keeps a's sign
#else
    return (a == 0) ? a : ((r < 0) ? step_down(q) : q);
#endif
}
MPH_HD double divk_ru(double a, double z, double zh, double zl) {
    const double q = fma(a, zh, a * zl);
    const double r = fma(-q, z, a);
#if defined(__HIP_DEVICE_COMPILE__)
    return (a == 0) ? a : hw_fma(r, 0x1p-64, q, true);
#else
    return (a == 0) ? a : ((r > 0) ? step_up(q) : q);
#endif
}
// er_md_rd / er_md_ru with z the interval evaluation's unit (frac in [1/2, 1)), through divk
MPH_HD Er er_mdk_rd(Er x, Er y, double zf, int ze, double zh, double zl) {
    return er_adjust(divk_rd(mul_rd(x.frac, y.frac), zf, zh, zl), x.exp + y.exp - ze);
}
MPH_HD Er er_mdk_ru(Er x, Er y, double zf, int ze, double zh, double zl) {
    return er_adjust(divk_ru(mul_ru(x.frac, y.frac), zf, zh, zl), x.exp + y.exp - ze);
}
// lo = er_mdk_rd(x0, y0), hi = er_mdk_ru(x1, y1) with the units (zf0, ze0) and (zf1, ze1)
MPH_HD void er_mdk_pair(Er x0, Er y0, Er x1, Er y1, double zf0, int ze0, double zh0, double zl0, double zf1, int ze1,
                        double zh1, double zl1, Er &lo, Er &hi) {
#if defined(__HIP_DEVICE_COMPILE__)
    double pl, ph;
    hw_mul_pair(x0.frac, y0.frac, x1.frac, y1.frac, pl, ph);
    const double ql = fma(pl, zh0, pl * zl0), qh = fma(ph, zh1, ph * zl1);
    const double rl = fma(-ql, zf0, pl), rh = fma(-qh, zf1, ph);
    double dl, dh;
    hw_fma_pair(rl, 0x1p-64, ql, rh, 0x1p-64, qh, dl, dh);  // the steps of divk_rd / divk_ru
    dl = (pl == 0) ? pl : dl;
    dh = (ph == 0) ? ph : dh;
#else
    const double dl = divk_rd(mul_rd(x0.frac, y0.frac), zf0, zh0, zl0);
    const double dh = divk_ru(mul_ru(x1.frac, y1.frac), zf1, zh1, zl1);
#endif
    lo = er_adjust(dl, x0.exp + y0.exp - ze0);
    hi = er_adjust(dh, x1.exp + y1.exp - ze1);
}

// ---------------------------------------------------------------------------
// Exact modular arithmetic for moduli m < 2^31 (values held as int).
// ---------------------------------------------------------------------------
// x in (-2m, 3m) -> [0, m)
MPH_HD double mod_fix(double r, double m) {
    r = (r < 0) ? r + m : r;
    r = (r < 0) ? r + m : r;
    r = (r >= m) ? r - m : r;
    r = (r >= m) ? r - m : r;
    return r;
}
// x in [-m, 2m) -> [0, m)
MPH_HD double mod_fix1(double r, double m) {
    r = (r < 0) ? r + m : r;
    return (r >= m) ? r - m : r;
}
// (a * b) mod m for 0 <= a, b < m < 2^31. With p = RN(a b) and pl = a b - p
// (exact, |pl| <= m^2 2^-53), RN(p RN(1/m)) differs from a b / m by less than
// m 2^-51 < 2^-20, so q = floor of it is within one of the true quotient and
// r = a b - q m (formed exactly: |p - q m| < 2^32) lies in [-m, 2m).
MPH_HD int mulmod(int a, int b, double m, double minv) {
    const double da = (double)a, db = (double)b;
    const double p = da * db;
    const double pl = fma(da, db, -p);
    const double q = floor(p * minv);
    double r = fma(-q, m, p) + pl;
    return (int)mod_fix1(r, m);
}
// v mod m for an integer-valued 0 <= v < 2^53 (q is again within one of v / m)
MPH_HD int modd(double v, double m, double minv) {
    const double q = floor(v * minv);
    return (int)mod_fix1(fma(-q, m, v), m);
}

}  // namespace mph
