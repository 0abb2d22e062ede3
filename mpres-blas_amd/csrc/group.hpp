// Multiple-precision numbers held by a group of L lanes of a wavefront.
//
// A number with N residues is spread over L consecutive lanes (L divides 64
// and N); lane l of the group holds residues l*R .. l*R+R-1 (R = N/L) and a
// copy of the scalar fields (sign, exponent, interval evaluation). L = N is
// one lane per modulus; L = 1 keeps a whole number in one lane. Fewer lanes
// per number means fewer replicated scalar instructions per number; more
// lanes means fewer residues (registers) per lane and parallel sums. The
// group's control flow is uniform (every branch depends on replicated
// scalars only), so cross-lane exchanges never read an inactive lane.
//
// Sums over the moduli (the pairwise psum of src/pairwise.cuh) are a
// balanced tree: the R values inside a lane first, then a butterfly over the
// L lanes. Each tree node is one directed-rounded addition and IEEE addition
// is commutative in every rounding mode, so every lane obtains exactly the
// reference's psum value. Mixed-radix conversion runs column by column (see
// rare.hpp).
//
// The multiply-accumulate steps of the reductions (fma_step, fma_d_step)
// stream the operands' residues from memory: the scalar part of the product
// and of the addition is decided first, then one pass over the residues
// forms product, alignment and sum, so only the accumulator's residues are
// held in registers. The rare cases (product needs rounding, sign straddles
// zero) fall back to the generic operations on fully loaded numbers.
#pragma once

#include <hip/hip_runtime.h>

#include "devconsts.hpp"
#include "numerics.hpp"
#include "ops.hpp"
#include "rare.hpp"

namespace mph {

template <int R>
struct MP {
    int d[R];
    int sign;
    int exp;
    Er ev0;  // eval[0]: lower bound of X/M
    Er ev1;  // eval[1]: upper bound of X/M
};

// mp_float_t arrays (array of structures, 4N+40 bytes per element)
template <int N>
struct AoSRef {
    static constexpr int STRIDE = 4 * N + 40;
    char *base;
    __device__ __forceinline__ int *digits(long long i) const { return (int *)(base + i * STRIDE); }
    __device__ __forceinline__ int *sign(long long i) const { return (int *)(base + i * STRIDE + 4 * N); }
    __device__ __forceinline__ int *exp(long long i) const { return (int *)(base + i * STRIDE + 4 * N + 4); }
    __device__ __forceinline__ ErMem *eval(long long i) const { return (ErMem *)(base + i * STRIDE + 4 * N + 8); }
};

// uniform access to the two layouts
template <int N>
__device__ __forceinline__ const int *digits_of(const SoA &v, long long i) { return v.d + i * N; }
template <int N>
__device__ __forceinline__ const int *digits_of(const AoSRef<N> &v, long long i) { return v.digits(i); }
template <int N>
__device__ __forceinline__ void scalars_of(const SoA &v, long long i, int &s, int &e, Er &e0, Er &e1) {
    s = v.s[i];
    e = v.e[i];
    const ErMem lo = v.ev[i], up = v.ev[i + v.len];
    e0 = Er{lo.frac, (int)lo.exp};
    e1 = Er{up.frac, (int)up.exp};
}
template <int N>
__device__ __forceinline__ void scalars_of(const AoSRef<N> &v, long long i, int &s, int &e, Er &e0, Er &e1) {
    s = *v.sign(i);
    e = *v.exp(i);
    const ErMem *p = v.eval(i);
    const ErMem lo = p[0], up = p[1];
    e0 = Er{lo.frac, (int)lo.exp};
    e1 = Er{up.frac, (int)up.exp};
}

// ---------------- scalar fields alone ----------------
// Sign, exponent and interval evaluation of a number. The split form of gemm
// runs a reduction's chain on these fields only and forms the residues
// separately (kernels.hpp).
struct Sc {
    int sign;
    int exp;
    Er ev0, ev1;
};

// exponent alignment shared by mp_add_common (src/arith/add.cuh:126-160) and the
// v1 vector/matrix add esi kernels (src/mpvector.cuh:275-299)
struct Align {
    int gamma, theta, nzx, nzy, sx, sy, fx, fy, ex, ey;
    Er x0, x1, y0, y1;
};
__device__ __forceinline__ Align align_sc(int mp_j, int xs, int xe, Er xv0, Er xv1, int ys, int ye, Er yv0, Er yv1) {
    Align a;
    const int dexp = xe - ye;
    a.gamma = dexp > 0 ? dexp : 0;
    a.theta = dexp < 0 ? -dexp : 0;
    a.nzx = (yv1.frac == 0) || (a.theta + yv1.exp < mp_j);
    a.nzy = (xv1.frac == 0) || (a.gamma + xv1.exp < mp_j);
    a.gamma *= a.nzy;
    a.theta *= a.nzx;
    a.ex = (xe - a.gamma) * a.nzx;
    a.ey = (ye - a.theta) * a.nzy;
    a.sx = xs * a.nzx;
    a.sy = ys * a.nzy;
    a.fx = (1 - 2 * a.sx) * a.nzx;
    a.fy = (1 - 2 * a.sy) * a.nzy;
    a.x0 = Er{xv0.frac * (double)a.fx, xv0.exp + a.gamma};
    a.x1 = Er{xv1.frac * (double)a.fx, xv1.exp + a.gamma};
    a.y0 = Er{yv0.frac * (double)a.fy, yv0.exp + a.theta};
    a.y1 = Er{yv1.frac * (double)a.fy, yv1.exp + a.theta};
    return a;
}
// The constants of the scalar steps, read once into registers by the kernels that run
// long scalar chains (so that no step waits on a reload from the constants block).
struct ScC {
    int mp_h, mp_j, uue, ule;
    double uuf, ulf, uu_rh, uu_rl, ul_rh, ul_rl;
};
__device__ __forceinline__ ScC sc_consts(const DevConsts *C) {
    return ScC{C->mp_h, C->mp_j, C->unit_upp.exp, C->unit_low.exp, C->unit_upp.frac, C->unit_low.frac,
               C->uu_rh, C->uu_rl, C->ul_rh, C->ul_rl};
}
__device__ __forceinline__ bool sc_needs_round(const ScC &K, const Er &ev1) {
    return ev1.frac != 0 && ev1.exp >= K.mp_h;
}
// scalar part of cuda::mp_mul (mp_mul_common, src/arith/mul.cuh:35-42)
__device__ __forceinline__ void sc_mul(const ScC &K, Sc &r, int xs, int xe, Er x0, Er x1, int ys, int ye, Er y0,
                                       Er y1) {
    r.exp = xe + ye;
    r.sign = xs ^ ys;
    er_mdk_pair(x0, y0, x1, y1, K.uuf, K.uue, K.uu_rh, K.uu_rl, K.ulf, K.ule, K.ul_rh, K.ul_rl, r.ev0, r.ev1);
}

// scalar part of cuda::mp_mul_d (src/arith/muld.cuh:59-67): y's significand (53-bit
// integer, 0 for y = 0) scales the bounds and its exponent adds to x's
__device__ __forceinline__ void sc_muld(Sc &r, int xs, int xe, Er x0, Er x1, double y) {
    const uint64_t u = d2u(y);
    const bool nz = y != 0;
    const double sig = nz ? (double)((u & ((((uint64_t)1) << 52) - 1)) | (((uint64_t)1) << 52)) : 0.0;
    r.exp = xe + ((int)((u >> 52) & 2047) - (nz ? 1075 : 0));
    r.sign = xs ^ (int)(u >> 63);
    double lo, hi;
    mul_rd_ru(x0.frac, sig, x1.frac, sig, lo, hi);
    r.ev0 = er_adjust(lo, x0.exp);
    r.ev1 = er_adjust(hi, x1.exp);
}

// acc = acc + p on the scalar fields, for chains whose every step is exact:
// false when the sum needs rounding, the sum's interval straddles zero, or the
// alignment drops a nonzero operand. The decisions are those of the lane-group
// step (Grp::step_tail) and add (Grp::add_try).
__device__ __forceinline__ bool sc_add_step(const ScC &C, Sc &acc, const Sc &p) {
    const Align a = align_sc(C.mp_j, acc.sign, acc.exp, acc.ev0, acc.ev1, p.sign, p.exp, p.ev0, p.ev1);
    if ((!a.nzx && acc.ev1.frac != 0) || (!a.nzy && p.ev1.frac != 0)) return false;
    Er e0, e1;
    er_add_pair(a.sx ? a.x1 : a.x0, a.sy ? a.y1 : a.y0, a.sx ? a.x0 : a.x1, a.sy ? a.y0 : a.y1, e0, e1);
    const int sign = e0.frac < 0;
    if (sign != (e1.frac < 0)) return false;
    acc.sign = sign;
    acc.exp = (a.ex == 0) ? a.ey : a.ex;
    if (sign) {
        acc.ev0 = Er{-e1.frac, e1.exp};
        acc.ev1 = Er{-1.0 * e0.frac, e0.exp};
    } else {
        acc.ev0 = e0;
        acc.ev1 = e1;
    }
    return !sc_needs_round(C, acc.ev1);
}
// acc = acc + x*y (y a double, mp_mul_d then mp_add) on the scalar fields, exact steps only
__device__ __forceinline__ bool sc_muld_step(const ScC &C, Sc &acc, int xs, int xe, Er x0, Er x1, double y) {
    Sc p;
    sc_muld(p, xs, xe, x0, x1, y);
    if (sc_needs_round(C, p.ev1)) return false;
    return sc_add_step(C, acc, p);
}

// acc = acc + x*y on the scalar fields, for chains whose every step is exact:
// false when the product or the sum needs rounding, the sum's interval
// straddles zero, or the alignment drops a nonzero operand. The decisions are
// those of the lane-group step (Grp::step_tail).
__device__ __forceinline__ bool sc_fma_step(const ScC &C, Sc &acc, int xs, int xe, Er x0, Er x1, int ys, int ye,
                                            Er y0, Er y1) {
    Sc p;
    sc_mul(C, p, xs, xe, x0, x1, ys, ye, y0, y1);
    if (sc_needs_round(C, p.ev1)) return false;
    return sc_add_step(C, acc, p);
}

template <int N, int L>
struct Grp {
    static_assert(N % L == 0, "L must divide N");
    static_assert((L & (L - 1)) == 0 && L <= 32, "L must be a power of two <= 32");
    static constexpr int R = N / L;
    static constexpr bool INL = (N <= 8);               // rare paths inlined and unrolled
    static constexpr bool CACHE = (L > 1 && R <= 8);    // per-lane moduli kept in registers
    static constexpr int CR = CACHE ? R : 1;

    Core<N, L> k;
    double cmd[CR], cmv[CR];
    int cmi[CR];
    ScC K;  // constants of the scalar steps, read once

    __device__ __forceinline__ explicit Grp(const DevConsts *c) : K(sc_consts(c)) {
        k.C = c;
        k.P2 = c->pow2;
        k.lane = (int)(__lane_id() & (L - 1));
        k.load();
        if constexpr (CACHE) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                cmd[r] = c->md[ix(r)];
                cmv[r] = c->minv[ix(r)];
                cmi[r] = c->m[ix(r)];
            }
        }
    }
    __device__ __forceinline__ int ix(int r) const { return k.lane * R + r; }
    __device__ __forceinline__ int lane() const { return k.lane; }
    __device__ __forceinline__ double md(int r) const {
        if constexpr (CACHE) return cmd[r];
        else return k.C->md[ix(r)];
    }
    __device__ __forceinline__ double mv(int r) const {
        if constexpr (CACHE) return cmv[r];
        else return k.C->minv[ix(r)];
    }
    __device__ __forceinline__ int mi(int r) const {
        if constexpr (CACHE) return cmi[r];
        else return k.C->m[ix(r)];
    }
    __device__ __forceinline__ int mmul(int a, int b, int r) const { return mulmod(a, b, md(r), mv(r)); }
    // Moduli read through a pointer the compiler must treat as new on every call
    // (HINT): keeps the fast kernels' multiply-accumulate loops from hoisting N
    // moduli and reciprocals into registers (they are re-read from the scalar
    // cache instead). The full kernels do not use the hint.
    template <bool HINT>
    __device__ __forceinline__ const DevConsts *fresh() const {
        const DevConsts *p = k.C;
#if defined(__HIP_DEVICE_COMPILE__)
        if constexpr (!CACHE && HINT) asm volatile("" : "+s"(p));
#endif
        return p;
    }
    __device__ __forceinline__ double md(int r, const DevConsts *c) const {
        if constexpr (CACHE) return cmd[r];
        else return c->md[ix(r)];
    }
    __device__ __forceinline__ double mv(int r, const DevConsts *c) const {
        if constexpr (CACHE) return cmv[r];
        else return c->minv[ix(r)];
    }
    __device__ __forceinline__ int mi(int r, const DevConsts *c) const {
        if constexpr (CACHE) return cmi[r];
        else return c->m[ix(r)];
    }
    __device__ __forceinline__ int mmul(int a, int b, int r, const DevConsts *c) const {
        return mulmod(a, b, md(r, c), mv(r, c));
    }
    __device__ __forceinline__ int prow(int g) const { return g > k.C->log2m ? k.C->log2m : g; }
    __device__ __forceinline__ Er unit_low() const { return Er{k.C->unit_low.frac, k.C->unit_low.exp}; }
    __device__ __forceinline__ Er unit_upp() const { return Er{k.C->unit_upp.frac, k.C->unit_upp.exp}; }

    // ---------------- rare paths ----------------
    __device__ __forceinline__ bool needs_round(const MP<R> &a) const {
        return a.ev1.frac != 0 && a.ev1.exp >= k.C->mp_h;
    }
    // mp_round with its callers' guard (src/arith/arith_utils.cuh:176-183, add.cuh:197)
    __device__ __forceinline__ void round_if_needed(MP<R> &a) const {
        if (needs_round(a)) {
            if constexpr (INL) {
                mp_round_body<N, L, true>(k, a.d, a.exp, a.ev0, a.ev1);
            } else {
                int t[R];
                int e = a.exp;
                Er e0 = a.ev0, e1 = a.ev1;
#pragma unroll
                for (int r = 0; r < R; r++) t[r] = a.d[r];
                ool_round<N, L>(k, t, &e, &e0, &e1);
#pragma unroll
                for (int r = 0; r < R; r++) a.d[r] = t[r];
                a.exp = e;
                a.ev0 = e0;
                a.ev1 = e1;
            }
        }
    }
    // full interval evaluation of a number's residues (rns_eval_compute, CUDA)
    __device__ __forceinline__ void eval_full(const int *d, Er &lo, Er &up) const {
        if constexpr (INL) {
            rns_eval<N, L, true, true>(k, d, lo, up, false);
        } else {
            int t[R];
            Er a, b;
#pragma unroll
            for (int r = 0; r < R; r++) t[r] = d[r];
            ool_eval_full<N, L>(k, t, &a, &b, false);
            lo = a;
            up = b;
        }
    }
    // sign_estimate, src/arith/arith_utils.cuh:199-207
    __device__ __forceinline__ int sign_estimate(const int *dx, const int *dy, int sx, int sy, int gamma, int theta,
                                                 int nzx, int nzy) const {
        int lx[R], ly[R];
        const int *rg = k.P2 + (size_t)prow(gamma) * N, *rt = k.P2 + (size_t)prow(theta) * N;
#pragma unroll
        for (int r = 0; r < R; r++) {
            lx[r] = nzx ? mmul(dx[r], rg[ix(r)], r) : 0;
            ly[r] = nzy ? mmul(dy[r], rt[ix(r)], r) : 0;
        }
        int cmp;
        if constexpr (INL)
            cmp = rns_mrc_compare<N, L, true>(k, lx, ly);
        else
            cmp = ool_mrc_compare<N, L>(k, lx, ly);
        return (cmp < 0 ? sy : sx) * (cmp != 0);
    }

    // rns_cmp with given interval evaluations, src/rns.cuh:1275-1289: the bounds decide
    // when they separate; otherwise equal residues give 0 and mixed-radix digits decide
    __device__ __forceinline__ int rns_cmp_ev(const int *lx, Er xl, Er xu, const int *ly, Er yl, Er yu) const {
        if (er_ucmp(xl, yu) > 0) return 1;
        if (er_ucmp(yl, xu) > 0) return -1;
        uint32_t diff = 0;
#pragma unroll
        for (int r = 0; r < R; r++) diff += lx[r] != ly[r];
        if (xlane_usum<L>(diff) == 0) return 0;
        if constexpr (INL)
            return rns_mrc_compare<N, L, true>(k, lx, ly);
        else
            return ool_mrc_compare<N, L>(k, lx, ly);
    }
    // cuda::mp_cmp_common (src/arith/cmp.cuh:79-114) and, with ABS, cuda::mp_cmp_abs_common
    // (src/arith/cmpabs.cuh:76-109): align to the larger exponent (dropping an operand that the
    // other's aligned evaluation dominates), compare magnitudes, then apply the signs
    template <bool ABS>
    __device__ __forceinline__ int cmp(const MP<R> &x, const MP<R> &y) const {
        const int dexp = x.exp - y.exp;
        int gamma = dexp > 0 ? dexp : 0, theta = dexp < 0 ? -dexp : 0;
        const int nzx = (y.ev1.frac == 0) || (theta + y.ev1.exp < k.C->mp_j);
        const int nzy = (x.ev1.frac == 0) || (gamma + x.ev1.exp < k.C->mp_j);
        gamma *= nzy;
        theta *= nzx;
        const Er xl{x.ev0.frac * (double)nzx, x.ev0.exp + gamma}, xu{x.ev1.frac * (double)nzx, x.ev1.exp + gamma};
        const Er yl{y.ev0.frac * (double)nzy, y.ev0.exp + theta}, yu{y.ev1.frac * (double)nzy, y.ev1.exp + theta};
        int lx[R], ly[R];
        const int *rg = k.P2 + (size_t)prow(gamma) * N, *rt = k.P2 + (size_t)prow(theta) * N;
#pragma unroll
        for (int r = 0; r < R; r++) {
            lx[r] = nzx ? mmul(x.d[r], rg[ix(r)], r) : 0;
            ly[r] = nzy ? mmul(y.d[r], rt[ix(r)], r) : 0;
        }
        const int c = rns_cmp_ev(lx, xl, xu, ly, yl, yu);
        if constexpr (ABS) {
            return c;
        } else {
            const int sx = x.sign, sy = y.sign;
            const bool gt = (sx == 0 && sy == 1) || (sx == 0 && sy == 0 && c == 1) || (sx == 1 && sy == 1 && c == -1);
            const bool lt = (sx == 1 && sy == 0) || (sx == 0 && sy == 0 && c == -1) || (sx == 1 && sy == 1 && c == 1);
            return gt ? 1 : lt ? -1 : 0;
        }
    }
    // cuda::rns_cmp(x, y) on bare residues, src/rns.cuh:1242-1260: both evaluations computed
    __device__ __forceinline__ int rns_cmp(const int *dx, const int *dy) const {
        Er xl, xu, yl, yu;
        eval_full(dx, xl, xu);
        eval_full(dy, yl, yu);
        return rns_cmp_ev(dx, xl, xu, dy, yl, yu);
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

    __device__ __forceinline__ Align align(int xs, int xe, Er xv0, Er xv1, int ys, int ye, Er yv0, Er yv1) const {
        return align_sc(k.C->mp_j, xs, xe, xv0, xv1, ys, ye, yv0, yv1);
    }
    __device__ __forceinline__ Align align(const MP<R> &x, const MP<R> &y) const {
        return align(x.sign, x.exp, x.ev0, x.ev1, y.sign, y.exp, y.ev0, y.ev1);
    }

    // residue of (2^gamma x fx + 2^theta y fy) mod m for one residue (at most one shift is nonzero)
    __device__ __forceinline__ int combine(int xd, int yd, const int *row, bool shx, bool sh, int fs, int fo,
                                          int sign, int r, const DevConsts *c) const {
        const int sv = shx ? xd : yd;
        const int ov = shx ? yd : xd;
        const int t = sh ? mmul(sv, row[ix(r)], r, c) : sv;
        const double v = mod_fix((double)fs * (double)t + (double)fo * (double)ov, md(r, c));
        const int vi = (int)v;
        return sign ? (vi != 0 ? mi(r, c) - vi : 0) : vi;
    }

    // cuda::mp_add (mp_add_common + rounding), src/arith/add.cuh:126-200
    __device__ __forceinline__ void add(MP<R> &res, const MP<R> &x, const MP<R> &y) const {
        const Align a = align(x, y);
        Er e0, e1;
        er_add_pair(a.sx ? a.x1 : a.x0, a.sy ? a.y1 : a.y0, a.sx ? a.x0 : a.x1, a.sy ? a.y0 : a.y1, e0, e1);
        int sign = e0.frac < 0;
        if (sign != (e1.frac < 0)) {
            sign = sign_estimate(x.d, y.d, a.sx, a.sy, a.gamma, a.theta, a.nzx, a.nzy);
            const Er u{k.C->unit_low.frac * (double)(1 - 2 * sign), k.C->unit_low.exp};
            if (sign)
                e1 = u;
            else
                e0 = u;
        }
        const bool shx = a.gamma > 0;
        const int sh = a.gamma + a.theta;
        const int *row = k.P2 + (size_t)prow(sh) * N;
        const int fs = shx ? a.fx : a.fy, fo = shx ? a.fy : a.fx;
#pragma unroll
        for (int r = 0; r < R; r++) res.d[r] = combine(x.d[r], y.d[r], row, shx, sh != 0, fs, fo, sign, r, k.C);
        finish_add(res, a, e0, e1, sign);
    }
    __device__ __forceinline__ void finish_add(MP<R> &res, const Align &a, Er e0, Er e1, int sign) const {
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

    // scalar part of cuda::mp_mul (mp_mul_common, src/arith/mul.cuh:35-42)
    __device__ __forceinline__ void mul_scalars(MP<R> &res, int xs, int xe, Er x0, Er x1, int ys, int ye, Er y0,
                                                Er y1) const {
        Sc p;
        sc_mul(K, p, xs, xe, x0, x1, ys, ye, y0, y1);
        res.sign = p.sign;
        res.exp = p.exp;
        res.ev0 = p.ev0;
        res.ev1 = p.ev1;
    }
    // cuda::mp_mul, src/arith/mul.cuh:53-74
    __device__ __forceinline__ void mul(MP<R> &res, const MP<R> &x, const MP<R> &y) const {
        mul_scalars(res, x.sign, x.exp, x.ev0, x.ev1, y.sign, y.exp, y.ev0, y.ev1);
#pragma unroll
        for (int r = 0; r < R; r++) res.d[r] = mmul(x.d[r], y.d[r], r);
        round_if_needed(res);
    }

    // scalar part of cuda::mp_mul_d (src/arith/muld.cuh:59-67); returns the significand of y
    __device__ __forceinline__ double muld_scalars(MP<R> &res, int xs, int xe, Er x0, Er x1, double y) const {
        const uint64_t u = d2u(y);
        const bool nz = y != 0;
        const double sig = nz ? (double)((u & ((((uint64_t)1) << 52) - 1)) | (((uint64_t)1) << 52)) : 0.0;
        res.exp = xe + ((int)((u >> 52) & 2047) - (nz ? 1075 : 0));
        res.sign = xs ^ (int)(u >> 63);
        res.ev0 = er_adjust(mul_rd(x0.frac, sig), x0.exp);
        res.ev1 = er_adjust(mul_ru(x1.frac, sig), x1.exp);
        return sig;
    }
    // cuda::mp_mul_d, src/arith/muld.cuh:59-73
    __device__ __forceinline__ void mul_d(MP<R> &res, const MP<R> &x, double y) const {
        const double sig = muld_scalars(res, x.sign, x.exp, x.ev0, x.ev1, y);
#pragma unroll
        for (int r = 0; r < R; r++) res.d[r] = mmul(x.d[r], modd(sig, md(r), mv(r)), r);
        round_if_needed(res);
    }

    // acc = acc + x[ix] * y[iy] (mp_mul then mp_add with the reference's rounding),
    // operands streamed from memory. FULL = false: returns false as soon as the
    // step needs a rare path (product or sum needs rounding, the sum's interval
    // straddles zero); acc is then unusable and the caller recomputes the chain.
    template <bool FULL, class VX, class VY>
    __device__ __forceinline__ bool fma_step(MP<R> &acc, const VX &vx, long long ixx, const VY &vy,
                                             long long iyy) const {
        int xs, xe, ys, ye;
        Er x0, x1, y0, y1;
        scalars_of<N>(vx, ixx, xs, xe, x0, x1);
        scalars_of<N>(vy, iyy, ys, ye, y0, y1);
        const int *dx = digits_of<N>(vx, ixx) + k.lane * R;
        const int *dy = digits_of<N>(vy, iyy) + k.lane * R;
        MP<R> p;
        mul_scalars(p, xs, xe, x0, x1, ys, ye, y0, y1);
        return step_tail<FULL>(acc, p, dx, dy, 0.0, false);
    }
    // as fma_step, with the lane's residues of x read from dxp (e.g. staged in LDS)
    template <bool FULL, class VX, class VY>
    __device__ __forceinline__ bool fma_step_d(MP<R> &acc, const VX &vx, long long ixx, const int *dxp, const VY &vy,
                                               long long iyy) const {
        int xs, xe, ys, ye;
        Er x0, x1, y0, y1;
        scalars_of<N>(vx, ixx, xs, xe, x0, x1);
        scalars_of<N>(vy, iyy, ys, ye, y0, y1);
        const int *dy = digits_of<N>(vy, iyy) + k.lane * R;
        MP<R> p;
        mul_scalars(p, xs, xe, x0, x1, ys, ye, y0, y1);
        return step_tail<FULL>(acc, p, dxp, dy, 0.0, false);
    }
    // acc = acc + x[ix] * y (mp_mul_d then mp_add), x streamed from memory
    template <bool FULL, class VX>
    __device__ __forceinline__ bool fma_d_step(MP<R> &acc, const VX &vx, long long ixx, double y) const {
        int xs, xe;
        Er x0, x1;
        scalars_of<N>(vx, ixx, xs, xe, x0, x1);
        const int *dx = digits_of<N>(vx, ixx) + k.lane * R;
        MP<R> p;
        const double sig = muld_scalars(p, xs, xe, x0, x1, y);
        return step_tail<FULL>(acc, p, dx, nullptr, sig, true);
    }
    // residue r of the product: x_r * y_r, or x_r * (sig mod m_r) for a double factor
    __device__ __forceinline__ int prod_digit(const int *dx, const int *dy, double sig, bool dbl, int r,
                                              const DevConsts *c) const {
        return dbl ? mmul(dx[r], modd(sig, md(r, c), mv(r, c)), r, c) : mmul(dx[r], dy[r], r, c);
    }
    template <bool FULL>
    __device__ __forceinline__ bool step_tail(MP<R> &acc, MP<R> &p, const int *dx, const int *dy, double sig,
                                              bool dbl) const {
        if (needs_round(p)) {
            if constexpr (!FULL) return false;
#pragma unroll
            for (int r = 0; r < R; r++) p.d[r] = prod_digit(dx, dy, sig, dbl, r, k.C);
            round_if_needed(p);
            add(acc, acc, p);
            return true;
        }
        const Align a = align(acc, p);
        Er e0, e1;
        er_add_pair(a.sx ? a.x1 : a.x0, a.sy ? a.y1 : a.y0, a.sx ? a.x0 : a.x1, a.sy ? a.y0 : a.y1, e0, e1);
        const int sign = e0.frac < 0;
        if (sign != (e1.frac < 0)) {
            if constexpr (!FULL) return false;
#pragma unroll
            for (int r = 0; r < R; r++) p.d[r] = prod_digit(dx, dy, sig, dbl, r, k.C);
            add(acc, acc, p);
            return true;
        }
        const bool shx = a.gamma > 0;
        const int sh = a.gamma + a.theta;
        const int *row = k.P2 + (size_t)prow(sh) * N;
        const int fs = shx ? a.fx : a.fy, fo = shx ? a.fy : a.fx;
        const DevConsts *c = fresh<!FULL>();
#pragma unroll
        for (int r = 0; r < R; r++)
            acc.d[r] = combine(acc.d[r], prod_digit(dx, dy, sig, dbl, r, c), row, shx, sh != 0, fs, fo, sign, r, c);
        acc.sign = sign;
        acc.exp = (a.ex == 0) ? a.ey : a.ex;
        if (sign) {
            acc.ev0 = Er{-e1.frac, e1.exp};
            acc.ev1 = Er{-1.0 * e0.frac, e0.exp};
        } else {
            acc.ev0 = e0;
            acc.ev1 = e1;
        }
        if (needs_round(acc)) {
            if constexpr (!FULL) return false;
            round_if_needed(acc);
        }
        return true;
    }
    // generic addition that may refuse the rare paths (see fma_step)
    template <bool FULL>
    __device__ __forceinline__ bool add_try(MP<R> &res, const MP<R> &x, const MP<R> &y) const {
        if constexpr (FULL) {
            add(res, x, y);
            return true;
        } else {
            const Align a = align(x, y);
            Er e0, e1;
            er_add_pair(a.sx ? a.x1 : a.x0, a.sy ? a.y1 : a.y0, a.sx ? a.x0 : a.x1, a.sy ? a.y0 : a.y1, e0, e1);
            const int sign = e0.frac < 0;
            if (sign != (e1.frac < 0)) return false;
            const bool shx = a.gamma > 0;
            const int sh = a.gamma + a.theta;
            const int *row = k.P2 + (size_t)prow(sh) * N;
            const int fs = shx ? a.fx : a.fy, fo = shx ? a.fy : a.fx;
            const DevConsts *c = fresh<!FULL>();
#pragma unroll
            for (int r = 0; r < R; r++) res.d[r] = combine(x.d[r], y.d[r], row, shx, sh != 0, fs, fo, sign, r, c);
            res.sign = sign;
            res.exp = (a.ex == 0) ? a.ey : a.ex;
            if (sign) {
                res.ev0 = Er{-e1.frac, e1.exp};
                res.ev1 = Er{-1.0 * e0.frac, e0.exp};
            } else {
                res.ev0 = e0;
                res.ev1 = e1;
            }
            return !needs_round(res);
        }
    }

    // v1 element-wise addition: esi + digits + round kernels, src/mpvector.cuh:234-472, 669-714
    // (no MRC on a straddling interval; residues via a double-reciprocal quotient)
    __device__ __forceinline__ void vadd(MP<R> &res, const MP<R> &x, const MP<R> &y) const {
        const Align a = align(x, y);
        Er e0, e1;
        er_add_pair(a.sx ? a.x1 : a.x0, a.sy ? a.y1 : a.y0, a.sx ? a.x0 : a.x1, a.sy ? a.y0 : a.y1, e0, e1);
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
        const int *rg = k.P2 + (size_t)prow(a.gamma) * N, *rt = k.P2 + (size_t)prow(a.theta) * N;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const long long rr = (long long)(a.fx * x.d[r]) * (long long)rg[ix(r)] +
                                 (long long)(a.fy * y.d[r]) * (long long)rt[ix(r)];
            const double q = (double)rr * mv(r);
            const int qi = (int)q;
            int v = (int)(rr - (long long)qi * (long long)mi(r));
            if (minus) {
                // cuda::mod_sub(m, v, m) = ((long)m - v) % m, C remainder semantics; |m - v| < 4m
                // needs 64 bits once moduli exceed 2^29
                const long long mm = mi(r);
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
            res.d[r] = v < 0 ? v + mi(r) : v;
        }
        if ((res.ev1.exp - k.C->mp_h + 1) * (res.ev1.frac != 0) > 0) round_if_needed(res);
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
        for (int r = 0; r < R; r++) res.d[r] = modd(sd, md(r), mv(r));
        // X < 2^53 <= m_0 ... m_{N-2} for every supported set, so mrd(X) = 0
        if constexpr (INL) {
            rns_eval<N, L, true, true>(k, res.d, res.ev0, res.ev1, true);
        } else {
            int t[R];
            Er a, b;
#pragma unroll
            for (int r = 0; r < R; r++) t[r] = res.d[r];
            ool_eval_full<N, L>(k, t, &a, &b, true);
            res.ev0 = a;
            res.ev1 = b;
        }
        round_if_needed(res);
    }
};

}  // namespace mph
