// Multiple-precision numbers held by a group of L lanes of a wavefront.
//
// A number with N residues is spread over L consecutive lanes (L divides 64
// and N); lane l of the group holds residues l*R .. l*R+R-1 (R = N/L) and a
// copy of the scalar fields (sign, exponent, interval evaluation). L = N is
// one lane per modulus; smaller L trades cross-lane traffic for fewer
// replicated scalar instructions. The group's control flow is uniform (every
// branch depends on replicated scalars only), so cross-lane exchanges never
// read an inactive lane.
//
// Sums over the moduli (the pairwise psum of src/pairwise.cuh) are a
// balanced tree: the R values inside a lane first, then a butterfly over the
// L lanes. Each tree node is one directed-rounded addition and IEEE addition
// is commutative in every rounding mode, so every lane obtains exactly the
// reference's psum value.
//
// Mixed-radix conversion (src/rns.cuh:570-630) runs column by column: step j
// broadcasts the finished digit j and every residue i > j applies its update
// (mr_i - mr_j) * m_j^-1 mod m_i, i.e. the reference's row loop reordered
// without changing any value: N steps instead of N(N-1)/2 serial products.
#pragma once

#include <hip/hip_runtime.h>

#include "devconsts.hpp"
#include "numerics.hpp"

namespace mph {

template <int R>
struct MP {
    int d[R];
    int sign;
    int exp;
    Er ev0;  // eval[0]: lower bound of X/M
    Er ev1;  // eval[1]: upper bound of X/M
};

constexpr int EVAL_MAX_ITER = 128;  // bound on the refinement loop; valid inputs need a handful

template <int N, int L>
struct Grp {
    static_assert(N % L == 0, "L must divide N");
    static_assert((L & (L - 1)) == 0 && L <= 32, "L must be a power of two <= 32");
    static constexpr int R = N / L;

    const DevConsts *__restrict__ C;
    const int *__restrict__ P2;
    int lane;
    double md[R], mv[R];
    int mi[R];

    __device__ __forceinline__ explicit Grp(const DevConsts *c) : C(c), P2(c->pow2) {
        lane = (int)(__lane_id() & (L - 1));
#pragma unroll
        for (int r = 0; r < R; r++) {
            md[r] = C->md[ix(r)];
            mv[r] = C->minv[ix(r)];
            mi[r] = C->m[ix(r)];
        }
    }
    __device__ __forceinline__ int ix(int r) const { return lane * R + r; }

    // ---------------- cross-lane primitives ----------------
    template <bool UP>
    __device__ __forceinline__ double psum(const double *v) const {
        double t[R];
#pragma unroll
        for (int r = 0; r < R; r++) t[r] = v[r];
#pragma unroll
        for (int w = 1; w < R; w <<= 1)
#pragma unroll
            for (int k = 0; k < R; k += 2 * w) t[k] = UP ? add_ru(t[k], t[k + w]) : add_rd(t[k], t[k + w]);
        double s = t[0];
#pragma unroll
        for (int off = 1; off < L; off <<= 1) {
            const double o = __shfl_xor(s, off, 64);
            s = UP ? add_ru(s, o) : add_rd(s, o);
        }
        return s;
    }
    __device__ __forceinline__ uint32_t usum(uint32_t v) const {
#pragma unroll
        for (int off = 1; off < L; off <<= 1) v += (uint32_t)__shfl_xor((int)v, off, 64);
        return v;
    }
    __device__ __forceinline__ int imax(int v) const {
#pragma unroll
        for (int off = 1; off < L; off <<= 1) {
            const int o = __shfl_xor(v, off, 64);
            v = o > v ? o : v;
        }
        return v;
    }
    // value of register r at group lane src (src, r compile-time after unrolling)
    __device__ __forceinline__ int bcast(int v, int src) const { return L == 1 ? v : __shfl(v, src, L); }

    // ---------------- modular helpers ----------------
    __device__ __forceinline__ int mmul(int a, int b, int r) const { return mulmod(a, b, md[r], mv[r]); }

    // ---------------- mixed-radix conversion ----------------
    __device__ __forceinline__ void mrc(int *mr, const int *x) const {
#pragma unroll
        for (int r = 0; r < R; r++) mr[r] = x[r];
#pragma unroll
        for (int j = 0; j < N - 1; j++) {
            const int v = bcast(mr[j % R], j / R);
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int i = ix(r);
                if (i > j) {
                    int t = mr[r] - v;
                    t = (mr[r] < v) ? t + mi[r] : t;
                    mr[r] = mmul(t, C->mrcinv[j][i], r);
                }
            }
        }
    }
    // most significant mixed-radix digit (src/rns.cuh:585-600)
    __device__ __forceinline__ int mrd(const int *x) const {
        int mr[R];
        mrc(mr, x);
        return bcast(mr[R - 1], L - 1);
    }
    // src/rns.cuh:624-630: 1 if x > y, -1 if x < y, 0 if equal (as integers < M)
    __device__ __forceinline__ int mrc_compare(const int *x, const int *y) const {
        int mx[R], my[R];
        mrc(mx, x);
        mrc(my, y);
        int key = -1;
#pragma unroll
        for (int r = 0; r < R; r++)
            if (mx[r] != my[r]) key = 2 * ix(r) + (mx[r] > my[r] ? 1 : 0);
        key = imax(key);
        return key < 0 ? 0 : ((key & 1) ? 1 : -1);
    }

    // ---------------- interval evaluation, src/rns.cuh:797-933 ----------------
    // FULL: rns_eval_compute (MRC on ambiguity); otherwise rns_eval_compute_fast.
    // known_small: X < m_0...m_{N-2} is known, so the most significant
    // mixed-radix digit is 0 and the MRC can be skipped without changing the result.
    template <bool FULL>
    __device__ __forceinline__ void eval(const int *x, Er &lo, Er &up, bool known_small = false) const {
        int s[R];
        double fl[R], fu[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            s[r] = mmul(x[r], C->w[ix(r)], r);
            fl[r] = mul_rd((double)s[r], C->rrd[ix(r)]);
            fu[r] = mul_ru((double)s[r], C->rru[ix(r)]);
        }
        double suml = psum<false>(fl);
        double sumu = psum<true>(fu);
        if (suml == 0 && sumu == 0) {
            lo = Er{0.0, 0};
            up = Er{0.0, 0};
            return;
        }
        const unsigned whl = (unsigned)suml, whu = (unsigned)sumu;
        suml = sub_rd(suml, (double)whl);
        sumu = sub_ru(sumu, (double)whu);
        const double acc = C->acc;
        if (!FULL) {
            if (sumu >= acc) {
                lo = er_set_d(suml);
                up = er_set_d(sumu);
                return;
            }
        } else {
            lo = er_set_d(suml);
            up = er_set_d(sumu);
            int mr = -1;
            if (whl != whu) mr = known_small ? 0 : mrd(x);
            if (mr > 0) {
                up = Er{C->inv_unit_upp.frac, C->inv_unit_upp.exp};
                return;
            }
            if (mr == 0) lo = Er{C->unit_low.frac, C->unit_low.exp};
            if (sumu >= acc) return;
        }
        int K = 0;
        const int ref = C->ref_factor, kmax = C->log2m;
        for (int it = 0; sumu < acc && it < EVAL_MAX_ITER; it++) {
            int k = -(ceil_log2_rn(sumu) + 1);
            k = k > ref ? k : ref;
            k = k < kmax ? k : kmax;
            const int *row = P2 + (size_t)k * N;
#pragma unroll
            for (int r = 0; r < R; r++) {
                s[r] = mmul(s[r], row[ix(r)], r);
                fu[r] = mul_ru((double)s[r], C->rru[ix(r)]);
            }
            sumu = psum<true>(fu);
            sumu = sub_ru(sumu, (double)(unsigned)sumu);
            K += k;
        }
#pragma unroll
        for (int r = 0; r < R; r++) fl[r] = mul_rd((double)s[r], C->rrd[ix(r)]);
        suml = psum<false>(fl);
        suml = sub_rd(suml, (double)(unsigned)suml);
        lo = er_set_d(suml);
        up = er_set_d(sumu);
        lo.exp -= K;
        up.exp -= K;
    }

    // ---------------- power-of-two scaling, src/rns.cuh:1061-1160 ----------------
    __device__ __forceinline__ void xw(int *c, const int *x) const {
#pragma unroll
        for (int r = 0; r < R; r++) c[r] = mmul(x[r], C->w[ix(r)], r);
    }
    __device__ __forceinline__ int rank_full(const int *x, const int *c) const {
        double fl[R], fu[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            fl[r] = mul_rd((double)c[r], C->rrd[ix(r)]);
            fu[r] = mul_ru((double)c[r], C->rru[ix(r)]);
        }
        const unsigned whl = (unsigned)psum<false>(fl), whu = (unsigned)psum<true>(fu);
        if (whl == whu) return (int)whl;
        return mrd(x) == 0 ? (int)whu : (int)whl;
    }
    __device__ __forceinline__ int rank_fast(const int *c) const {
        double fu[R];
#pragma unroll
        for (int r = 0; r < R; r++) fu[r] = mul_ru((double)c[r], C->rru[ix(r)]);
        return (int)psum<true>(fu);
    }
    // y = (x - (X mod 2^j)) * 2^-j, given the rank k of x (make_scaling_step, :1104-1124)
    __device__ __forceinline__ void scale_step(int *y, int k, int j, const int *x, const int *c) const {
        const uint32_t mask = (j >= 32) ? 0xffffffffu : ((1u << j) - 1u);
        uint32_t part = 0;
#pragma unroll
        for (int r = 0; r < R; r++) part += ((uint32_t)C->h[j - 1][ix(r)] * (uint32_t)c[r]) & mask;
        uint32_t res = usum(part);
        res = (res - (uint32_t)k * (uint32_t)C->mpow2[j - 1]) & mask;
#pragma unroll
        for (int r = 0; r < R; r++) {
            int mult = modd((double)res, md[r], mv[r]);
            int t = x[r] - mult;
            t = t < 0 ? t + mi[r] : t;
            y[r] = mmul(t, C->pow2inv[j - 1][ix(r)], r);
        }
    }
    __device__ __forceinline__ void scale2pow(int *x, unsigned D) const {
        int c[R];
        int t = (int)(D / P2T);
        if (t > 0) {
            xw(c, x);
            const int k = rank_full(x, c);
            scale_step(x, k, P2T, x, c);
            t -= 1;
        }
        while (t > 0) {
            xw(c, x);
            const int k = rank_fast(c);
            scale_step(x, k, P2T, x, c);
            t -= 1;
        }
        const unsigned d = D % P2T;
        if (d > 0) {
            xw(c, x);
            const int k = (d < D) ? rank_full(x, c) : rank_fast(c);
            scale_step(x, k, (int)d, x, c);
        }
    }

    // mp_round with its callers' guard (src/arith/arith_utils.cuh:176-183, add.cuh:197)
    __device__ __forceinline__ void round_if_needed(MP<R> &a) const {
        if (a.ev1.frac != 0 && a.ev1.exp >= C->mp_h) {
            const int bits = a.ev1.exp - C->mp_h + 1;
            a.exp += bits;
            scale2pow(a.d, (unsigned)bits);
            eval<false>(a.d, a.ev0, a.ev1);
        }
    }

    // ---------------- mp_float_t arithmetic ----------------
    __device__ __forceinline__ void zero(MP<R> &a) const {
#pragma unroll
        for (int r = 0; r < R; r++) a.d[r] = 0;
        a.sign = 0;
        a.exp = 0;
        a.ev0 = Er{0.0, 0};
        a.ev1 = Er{0.0, 0};
    }
    __device__ __forceinline__ int prow(int g) const { return g > C->log2m ? C->log2m : g; }

    // sign_estimate, src/arith/arith_utils.cuh:199-207
    __device__ __forceinline__ int sign_estimate(const int *dx, const int *dy, int sx, int sy, int gamma, int theta,
                                                 int nzx, int nzy) const {
        int lx[R], ly[R];
        const int *rg = P2 + (size_t)prow(gamma) * N, *rt = P2 + (size_t)prow(theta) * N;
#pragma unroll
        for (int r = 0; r < R; r++) {
            lx[r] = nzx ? mmul(dx[r], rg[ix(r)], r) : 0;
            ly[r] = nzy ? mmul(dy[r], rt[ix(r)], r) : 0;
        }
        const int cmp = mrc_compare(lx, ly);
        return (cmp < 0 ? sy : sx) * (cmp != 0);
    }

    // exponent alignment shared by mp_add_common (src/arith/add.cuh:126-160) and the
    // v1 vector/matrix add esi kernels (src/mpvector.cuh:275-299)
    struct Align {
        int gamma, theta, nzx, nzy, sx, sy, fx, fy, ex, ey;
        Er x0, x1, y0, y1;
    };
    __device__ __forceinline__ Align align(const MP<R> &x, const MP<R> &y) const {
        Align a;
        const int dexp = x.exp - y.exp;
        a.gamma = dexp > 0 ? dexp : 0;
        a.theta = dexp < 0 ? -dexp : 0;
        const int mpj = C->mp_j;
        a.nzx = (y.ev1.frac == 0) || (a.theta + y.ev1.exp < mpj);
        a.nzy = (x.ev1.frac == 0) || (a.gamma + x.ev1.exp < mpj);
        a.gamma *= a.nzy;
        a.theta *= a.nzx;
        a.ex = (x.exp - a.gamma) * a.nzx;
        a.ey = (y.exp - a.theta) * a.nzy;
        a.sx = x.sign * a.nzx;
        a.sy = y.sign * a.nzy;
        a.fx = (1 - 2 * a.sx) * a.nzx;
        a.fy = (1 - 2 * a.sy) * a.nzy;
        a.x0 = Er{x.ev0.frac * (double)a.fx, x.ev0.exp + a.gamma};
        a.x1 = Er{x.ev1.frac * (double)a.fx, x.ev1.exp + a.gamma};
        a.y0 = Er{y.ev0.frac * (double)a.fy, y.ev0.exp + a.theta};
        a.y1 = Er{y.ev1.frac * (double)a.fy, y.ev1.exp + a.theta};
        return a;
    }

    // cuda::mp_add (mp_add_common + rounding), src/arith/add.cuh:126-200
    __device__ __forceinline__ void add(MP<R> &res, const MP<R> &x, const MP<R> &y) const {
        const Align a = align(x, y);
        Er e0 = er_add_dir(a.sx ? a.x1 : a.x0, a.sy ? a.y1 : a.y0, false);
        Er e1 = er_add_dir(a.sx ? a.x0 : a.x1, a.sy ? a.y0 : a.y1, true);
        int sign = e0.frac < 0;
        if (sign != (e1.frac < 0)) {
            sign = sign_estimate(x.d, y.d, a.sx, a.sy, a.gamma, a.theta, a.nzx, a.nzy);
            const Er u{C->unit_low.frac * (double)(1 - 2 * sign), C->unit_low.exp};
            if (sign)
                e1 = u;
            else
                e0 = u;
        }
        // residues: (2^gamma * x * fx + 2^theta * y * fy) mod m; at most one of gamma, theta is nonzero
        const bool shx = a.gamma > 0;
        const int *row = P2 + (size_t)prow(a.gamma + a.theta) * N;
        const int fs = shx ? a.fx : a.fy, fo = shx ? a.fy : a.fx;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int sv = shx ? x.d[r] : y.d[r];
            const int ov = shx ? y.d[r] : x.d[r];
            const double t = (double)mmul(sv, row[ix(r)], r);
            double v = (double)fs * t + (double)fo * (double)ov;
            v = mod_fix(v, md[r]);
            int vi = (int)v;
            res.d[r] = sign ? (vi != 0 ? mi[r] - vi : 0) : vi;
        }
        res.sign = sign;
        res.exp = (a.ex == 0) ? a.ey : a.ex;
        if (sign) {
            res.ev0 = Er{-e1.frac, e1.exp};
            res.ev1 = Er{-1.0 * e0.frac, e0.exp};
        } else {
            res.ev0 = e0;
            res.ev1 = e1;
        }
        round_if_needed(res);
    }

    // cuda::mp_mul (mp_mul_common + rounding), src/arith/mul.cuh:53-74
    __device__ __forceinline__ void mul(MP<R> &res, const MP<R> &x, const MP<R> &y) const {
        res.exp = x.exp + y.exp;
        res.sign = x.sign ^ y.sign;
        const Er ul{C->unit_low.frac, C->unit_low.exp}, uu{C->unit_upp.frac, C->unit_upp.exp};
        res.ev0 = er_md_rd(x.ev0, y.ev0, uu);
        res.ev1 = er_md_ru(x.ev1, y.ev1, ul);
#pragma unroll
        for (int r = 0; r < R; r++) res.d[r] = mmul(x.d[r], y.d[r], r);
        round_if_needed(res);
    }

    // cuda::mp_mul_d, src/arith/muld.cuh:59-73
    __device__ __forceinline__ void mul_d(MP<R> &res, const MP<R> &x, double y) const {
        const uint64_t u = d2u(y);
        const bool nz = y != 0;
        const double sig = nz ? (double)((u & ((((uint64_t)1) << 52) - 1)) | (((uint64_t)1) << 52)) : 0.0;
        res.exp = x.exp + ((int)((u >> 52) & 2047) - (nz ? 1075 : 0));
        res.sign = x.sign ^ (int)(u >> 63);
        res.ev0 = er_adjust(mul_rd(x.ev0.frac, sig), x.ev0.exp);
        res.ev1 = er_adjust(mul_ru(x.ev1.frac, sig), x.ev1.exp);
#pragma unroll
        for (int r = 0; r < R; r++) res.d[r] = mmul(x.d[r], modd(sig, md[r], mv[r]), r);
        round_if_needed(res);
    }

    // v1 element-wise addition: esi + digits + round kernels, src/mpvector.cuh:234-472, 669-714
    // (no MRC on a straddling interval; residues via a double-reciprocal quotient)
    __device__ __forceinline__ void vadd(MP<R> &res, const MP<R> &x, const MP<R> &y) const {
        const Align a = align(x, y);
        Er e0 = er_add_dir(a.sx ? a.x1 : a.x0, a.sy ? a.y1 : a.y0, false);
        Er e1 = er_add_dir(a.sx ? a.x0 : a.x1, a.sy ? a.y0 : a.y1, true);
        const int minus = e0.frac < 0 && e1.frac < 0;
        res.sign = minus;
        res.exp = (a.ex == 0) ? a.ey : a.ex;
        if (minus) {
            res.ev0 = Er{-e1.frac, e1.exp};
            res.ev1 = Er{-e0.frac, e0.exp};
        } else {
            res.ev0 = e0;
            res.ev1 = e1;
        }
        const int *rg = P2 + (size_t)prow(a.gamma) * N, *rt = P2 + (size_t)prow(a.theta) * N;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const long long rr = (long long)(a.fx * x.d[r]) * (long long)rg[ix(r)] +
                                 (long long)(a.fy * y.d[r]) * (long long)rt[ix(r)];
            const double q = (double)rr * mv[r];
            const int qi = (int)q;
            int v = (int)(rr - (long long)qi * (long long)mi[r]);
            if (minus) {
                // cuda::mod_sub(m, v, m) = ((long)m - v) % m, C remainder semantics; |m - v| < 4m
                // needs 64 bits once moduli exceed 2^29
                const long long mm = mi[r];
                long long t = mm - (long long)v;
                if (t >= 0) {
                    t = t >= mm ? t - mm : t;
                    t = t >= mm ? t - mm : t;
                    t = t >= mm ? t - mm : t;
                } else {
                    t = t <= -mm ? t + mm : t;
                    t = t <= -mm ? t + mm : t;
                    t = t <= -mm ? t + mm : t;
                }
                v = (int)t;
            }
            res.d[r] = v < 0 ? v + mi[r] : v;
        }
        const int bits = (res.ev1.exp - C->mp_h + 1) * (res.ev1.frac != 0);
        if (bits > 0) {
            res.exp += bits;
            scale2pow(res.d, (unsigned)bits);
            eval<false>(res.d, res.ev0, res.ev1);
        }
    }

    // mp_set_d, src/arith/assign.cuh:54-81, with the CUDA interval evaluation
    __device__ __forceinline__ void set_d(MP<R> &res, double x) const {
        const uint64_t u = d2u(x);
        res.sign = (int)(u >> 63);
        int e = (int)((u >> 52) & 2047);
        uint64_t sig = (u & ((((uint64_t)1) << 52) - 1)) | (e != 0 ? (((uint64_t)1) << 52) : 0);
        e = e - 1023 - 52;
        if (sig != 0) {
            const int tz = __builtin_ctzll(sig);
            sig >>= tz;
            e += tz;
        } else {
            e = 0;
        }
        res.exp = e;
        const double sd = (double)sig;
#pragma unroll
        for (int r = 0; r < R; r++) res.d[r] = modd(sd, md[r], mv[r]);
        // X < 2^53 <= m_0 ... m_{N-2} for every supported set, so mrd(X) = 0
        eval<true>(res.d, res.ev0, res.ev1, true);
        round_if_needed(res);
    }
};

}  // namespace mph
