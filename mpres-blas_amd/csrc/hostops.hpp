// Device versions of the operations the reference runs on the host with MPFR: the CUDA
// mp_div, mp_sqrt and mp_dsqrt copy their operands to the host and call the CPU functions
// (src/arith/div.cuh:31-65, src/arith/sqrt.cuh:33-89). Each result is defined by MPFR:
//   mp_get_mpfr  X * 2^exp rounded to MP_PRECISION bits, to nearest (assign.cuh:185-211)
//   mpfr_div / mpfr_sqrt at MP_PRECISION bits, to nearest
//   mp_set_mpfr  the significand without trailing zeros, its residues, and the CPU
//                rns_eval_compute of those residues (assign.cuh:86-127)
//   mpfr_get_d   to nearest
// All of these are exact integer operations (bigint.hpp) followed by the CPU interval
// evaluation, restated here with the CPU path's arithmetic: emulated directed rounding
// RN(c) -/+ (phi |c| + eta) (src/dinterval.cuh:55-165), products mod m through a double
// quotient (src/modular.cuh:90-95, 218-222), the zero test on the residues and division
// in the refinement (src/rns.cuh:645-716). One thread per number.
#pragma once

#include <hip/hip_runtime.h>

#include "bigint.hpp"
#include "devconsts.hpp"
#include "numerics.hpp"

namespace mph {

// ---- the CPU path's arithmetic (src/dinterval.cuh, src/pairwise.cuh:402-479) ----
constexpr double E_PHI1 = 0x1.0000000000001p-53;  // DBL_EPS (1 + 2 DBL_EPS), src/bitwise.cuh:51-61
constexpr double E_ETA = 0x1p-1074;               // DBL_ETA
constexpr int E_EVAL_MAX_ITER = 128;               // bound on the refinement loop (valid inputs need a few)
__device__ __forceinline__ double e_widen(double c, bool up) {
    const double e = E_PHI1 * fabs(c) + E_ETA;  // unfused (fmac without EMPLOY_STD_FMA)
    return up ? c + e : c - e;
}
template <int N>
__device__ __forceinline__ double e_psum(const double *v, bool up) {
    double t[N];
#pragma unroll
    for (int i = 0; i < N; i++) t[i] = v[i];
#pragma unroll
    for (int w = 1; w < N; w <<= 1)
#pragma unroll
        for (int k = 0; k < N; k += 2 * w) t[k] = e_widen(t[k] + t[k + w], up);
    return e_widen(t[0] + 0.0, up);  // the root addition of 0 of psum(2N)<N>
}
// (x * y) mod m through the double quotient: may return values just below 0 or at or
// above m, as the reference's CPU mod_mul does
__device__ __forceinline__ int e_mod_mul(int x, int y, int m, double d) {
    const long long r = (long long)x * (long long)y;
    const int q = (int)((double)r * d);
    return (int)(r - (long long)q * (long long)m);
}
// last mixed-radix digit, src/rns.cuh:495-520 (exact)
template <int N>
__device__ __forceinline__ int e_mrd(const DevConsts *__restrict__ C, const int *x) {
    int mr[N];
#pragma unroll
    for (int i = 0; i < N; i++) mr[i] = x[i];
#pragma unroll
    for (int j = 0; j < N - 1; j++)
#pragma unroll
        for (int i = j + 1; i < N; i++) {
            int t = mr[i] - mr[j];
            t = t < 0 ? t + C->m[i] : t;
            mr[i] = mulmod(t, C->mrcinv[j][i], C->md[i], C->minv[i]);
        }
    return mr[N - 1];
}
// rns_eval_compute of the CPU path, src/rns.cuh:645-716
template <int N>
__device__ void cpu_eval(const DevConsts *__restrict__ C, const int *x, Er &low, Er &upp) {
    bool zero = true;
#pragma unroll
    for (int i = 0; i < N; i++) zero = zero && x[i] == 0;
    if (zero) {
        low = Er{0.0, 0};
        upp = Er{0.0, 0};
        return;
    }
    int s[N];
    double fl[N], fu[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        s[i] = e_mod_mul(x[i], C->w[i], C->m[i], C->minv[i]);
        fl[i] = e_widen((double)s[i] * C->rrd[i], false);
        fu[i] = e_widen((double)s[i] * C->rru[i], true);
    }
    double suml = e_psum<N>(fl, false), sumu = e_psum<N>(fu, true);
    const unsigned whl = (unsigned)suml, whu = (unsigned)sumu;
    suml = suml - whl;
    sumu = sumu - whu;
    low = er_set_d(suml);
    upp = er_set_d(sumu);
    int mr = -1;
    if (whl != whu) mr = e_mrd<N>(C, x);
    if (mr > 0) {
        upp = Er{C->inv_unit_upp.frac, C->inv_unit_upp.exp};
        return;
    }
    if (mr == 0) low = Er{C->unit_low.frac, C->unit_low.exp};
    if (sumu >= C->acc) return;
    int K = 0;
    for (int it = 0; it < E_EVAL_MAX_ITER && sumu < C->acc; it++) {
        int k = -(ceil_log2_rn(sumu) + 1);
        k = k > C->ref_factor ? k : C->ref_factor;
#pragma unroll
        for (int i = 0; i < N; i++) {
            s[i] = e_mod_mul(s[i], C->pow2[(size_t)k * N + i], C->m[i], C->minv[i]);
            fu[i] = e_widen((double)s[i] / C->md[i], true);
        }
        sumu = e_psum<N>(fu, true);
        sumu -= (unsigned)sumu;
        K += k;
    }
#pragma unroll
    for (int i = 0; i < N; i++) fl[i] = e_widen((double)s[i] / C->md[i], false);
    suml = e_psum<N>(fl, false);
    suml -= (unsigned)suml;
    low = er_set_d(suml);
    upp = er_set_d(sumu);
    low.exp -= K;
    upp.exp -= K;
}

// ---- exact values ----
// limbs of the integers: X < M < 2^(31 N), and the intermediate values of the division
// and square root at MP_PRECISION < 31 N / 2 bits
template <int N>
constexpr int big_limbs() {
    return (31 * N + 31) / 32 + 2;
}
template <int N>
using BigN = Big<big_limbs<N>()>;

// X of a number's residues (0 <= X < M), mixed-radix conversion then Horner's rule
template <int N>
__device__ BigN<N> residues_to_big(const DevConsts *__restrict__ C, const int *d) {
    int a[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int m = C->m[i];
        const int v = d[i] % m;
        a[i] = v < 0 ? v + m : v;
    }
#pragma unroll
    for (int j = 0; j < N - 1; j++)
#pragma unroll
        for (int i = j + 1; i < N; i++) {
            int t = a[i] - a[j];
            t = t < 0 ? t + C->m[i] : t;
            a[i] = mulmod(t, C->mrcinv[j][i], C->md[i], C->minv[i]);
        }
    BigN<N> X;
    big_zero(X);
    X.l[0] = (uint32_t)a[N - 1];
    for (int i = N - 2; i >= 0; i--) big_mul_add_u32(X, (uint32_t)C->m[i], (uint32_t)a[i]);
    return X;
}

// mp_set_mpfr of (-1)^sign Q 2^e with Q at most MP_PRECISION + 1 bits (assign.cuh:86-127):
// Q without trailing zeros, its residues, the CPU evaluation. A zero keeps the exponent
// that mpfr_get_str's string gives it: -1 for +0, 0 for -0.
template <int N>
__device__ void big_to_mp(const DevConsts *__restrict__ C, BigN<N> Q, int e, int sign, int *digits, int &osign,
                          int &oexp, Er &lo, Er &up) {
    osign = sign;
    if (big_is_zero(Q)) {
        oexp = sign ? 0 : -1;
#pragma unroll
        for (int i = 0; i < N; i++) digits[i] = 0;
        lo = Er{0.0, 0};
        up = Er{0.0, 0};
        return;
    }
    const int tz = big_ctz(Q);
    Q = big_shr(Q, tz);
    oexp = e + tz;
#pragma unroll
    for (int i = 0; i < N; i++) digits[i] = (int)big_mod_u32(Q, (uint32_t)C->m[i]);
    cpu_eval<N>(C, digits, lo, up);
}

// mp_get_mpfr into MP_PRECISION bits: X rounded to nearest, the exponent adjusted
template <int N>
__device__ BigN<N> mp_to_big(const DevConsts *__restrict__ C, const int *d, int exp, int &e) {
    BigN<N> X = residues_to_big<N>(C, d);
    e = exp;
    big_round(X, e, C->mp_precision);
    return X;
}

// mpfr_div to nearest at P = MP_PRECISION bits of a 2^ea / b 2^eb (b != 0): q = P + 2
// quotient bits of the operands aligned to one length, rounded to P bits with the
// remainder as sticky bit
template <int N>
__device__ BigN<N> big_quotient(BigN<N> a, int ea, BigN<N> b, int eb, int P, int &e) {
    const int la = big_bitlen(a), lb = big_bitlen(b);
    const int L = la > lb ? la : lb;
    a = big_shl(a, L - la);
    b = big_shl(b, L - lb);
    bool inexact;
    BigN<N> q = big_div_bits(a, b, P + 2, inexact);  // floor(a/b 2^(P+1))
    const int bq = big_bitlen(q), sh = bq - P;       // 1 or 2
    const int half = big_bit(q, sh - 1);
    const bool sticky = inexact || big_any_below(q, sh - 1);
    q = big_shr(q, sh);
    if (half && (sticky || (q.l[0] & 1u))) big_add_u32(q, 1u);
    e = sh - (P + 1) + ea - eb + la - lb;
    return q;
}

// mpfr_sqrt to nearest at P bits of a 2^ea (a != 0): the exponent made even and the
// radicand widened to 2P + 2 or 2P + 3 bits, its integer square root (P + 1 or P + 2
// bits) rounded to P bits with the remainder as sticky bit
template <int N>
__device__ Big<2 * big_limbs<N>()> big_sqrt_round(const BigN<N> &a, int ea, int P, int &e) {
    constexpr int W2 = 2 * big_limbs<N>();
    Big<W2> x;
    big_zero(x);
#pragma unroll
    for (int i = 0; i < big_limbs<N>(); i++) x.l[i] = a.l[i];
    int la = big_bitlen(x);
    int sh = 2 * P + 2 - la;                // widen to at least 2P + 2 bits
    sh = sh < 0 ? 0 : sh;
    if ((ea - sh) & 1) sh++;                // keep the exponent even
    x = big_shl(x, sh);
    int ex = ea - sh;
    bool inexact;
    Big<W2> s = big_isqrt(x, inexact);
    const int bs = big_bitlen(s), d = bs - P;
    int es = ex / 2;
    if (d > 0) {
        const int half = big_bit(s, d - 1);
        const bool sticky = inexact || big_any_below(s, d - 1);
        s = big_shr(s, d);
        if (half && (sticky || (s.l[0] & 1u))) big_add_u32(s, 1u);
        es += d;
    }
    e = es;
    return s;
}

// mpfr_get_d to nearest of Q 2^e (Q up to W limbs): one rounding, to 53 significant bits
// for a normal result and to the bits above 2^-1074 for a subnormal one (ties to even;
// infinity on overflow)
template <int W>
__device__ double big_to_double(Big<W> Q, int e, int sign) {
    if (big_is_zero(Q)) return sign ? -0.0 : 0.0;
    const int E = e + big_bitlen(Q) - 1;  // exponent of the leading bit
    const int p = E >= -1022 ? 53 : E + 1075;
    double v;
    if (p >= 1) {
        big_round(Q, e, p);
        v = ldexp((double)(((uint64_t)Q.l[1] << 32) | Q.l[0]), e);  // exact
    } else if (p == 0) {
        // Q 2^e in [2^-1075, 2^-1074): half the least subnormal exactly when Q is a power of two
        v = (big_ctz(Q) == big_bitlen(Q) - 1) ? 0.0 : 0x1p-1074;
    } else {
        v = 0.0;
    }
    return sign ? -v : v;
}

}  // namespace mph
