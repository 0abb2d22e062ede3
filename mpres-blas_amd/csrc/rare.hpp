// The non-positional RNS algorithms that only run when a result needs
// rounding, an interval straddles zero, or a conversion needs refinement:
// mixed-radix conversion and comparison, interval evaluation, rank, and
// power-of-two scaling (src/rns.cuh:570-630, 797-933, 1061-1160).
//
// They are written once against a small core (constants pointer and lane
// index) and per-lane residue arrays, and instantiated two ways:
//   U = true : loops unrolled, arrays in registers, inlined into the caller
//              (8 moduli: rounding happens on almost every 106-bit addition);
//   U = false: rolled loops, arrays in private memory, called out of line
//              (16 and 32 moduli: rounding is rare, and keeping these paths
//              out of the hot loops keeps their register footprint small).
#pragma once

#include <hip/hip_runtime.h>

#include "devconsts.hpp"
#include "numerics.hpp"

namespace mph {

template <int N, int L>
struct Core {
    static constexpr int R = N / L;
    // One residue per lane on the inlined paths (8 moduli on 8 lanes, where rounding runs on
    // almost every step): the lane's constants are read once, by load(), and held in
    // registers, so that no step waits on a reload of them. Otherwise they are read where used.
    static constexpr bool KC = (N <= 8 && R == 1);
    static constexpr int KN = KC ? N - 1 : 1;
    const DevConsts *__restrict__ C;
    const int *__restrict__ P2;
    int lane;
    double kmd, kmv, krd, kru;
    int kmi, kw, kmrc[KN];
    __device__ __forceinline__ void load() {
        if constexpr (KC) {
            kmd = C->md[lane];
            kmv = C->minv[lane];
            krd = C->rrd[lane];
            kru = C->rru[lane];
            kmi = C->m[lane];
            kw = C->w[lane];
#pragma unroll
            for (int j = 0; j < KN; j++) kmrc[j] = C->mrcinv[j][lane];
        }
    }
    __device__ __forceinline__ int ix(int r) const { return lane * R + r; }
    __device__ __forceinline__ double md(int r) const {
        if constexpr (KC) return kmd;
        else return C->md[ix(r)];
    }
    __device__ __forceinline__ double mv(int r) const {
        if constexpr (KC) return kmv;
        else return C->minv[ix(r)];
    }
    __device__ __forceinline__ int mi(int r) const {
        if constexpr (KC) return kmi;
        else return C->m[ix(r)];
    }
    __device__ __forceinline__ int w(int r) const {
        if constexpr (KC) return kw;
        else return C->w[ix(r)];
    }
    __device__ __forceinline__ double rrd(int r) const {
        if constexpr (KC) return krd;
        else return C->rrd[ix(r)];
    }
    __device__ __forceinline__ double rru(int r) const {
        if constexpr (KC) return kru;
        else return C->rru[ix(r)];
    }
    // m_j^-1 mod m_i for j < i = ix(r) (the mixed-radix conversion's step j)
    __device__ __forceinline__ int mrc(int j, int r) const {
        if constexpr (KC) return kmrc[j];
        else return C->mrcinv[j][ix(r)];
    }
};

template <int COUNT, bool U, typename F>
__device__ __forceinline__ void loop(F &&f) {
    if constexpr (U) {
#pragma unroll
        for (int j = 0; j < COUNT; j++) f(j);
    } else {
#pragma nounroll
        for (int j = 0; j < COUNT; j++) f(j);
    }
}

// ---------------- cross-lane sums (balanced tree, see group.hpp) ----------------
// The value of lane (lane ^ OFF) at one level of a butterfly whose earlier levels made
// every lane of each block of OFF lanes equal: OFF = 1, 2 by quad permutations, 4 and 8 by
// the mirror within 8 and 16 lanes (it reaches the other block of OFF lanes, whose lanes
// all hold its value), all by DPP in registers; 16 and more through a shuffle.
template <int OFF>
__device__ __forceinline__ int xchg(int v) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (OFF == 1) return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);       // quad_perm 1,0,3,2
    else if constexpr (OFF == 2) return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);  // quad_perm 2,3,0,1
    else if constexpr (OFF == 4) return __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false); // row_half_mirror
    else if constexpr (OFF == 8) return __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false); // row_mirror
    else return __shfl_xor(v, OFF, 64);
#else
    return __shfl_xor(v, OFF, 64);
#endif
}
template <int OFF>
__device__ __forceinline__ double xchg_d(double v) {
    const uint64_t u = d2u(v);
    const uint32_t lo = (uint32_t)xchg<OFF>((int)(uint32_t)u), hi = (uint32_t)xchg<OFF>((int)(uint32_t)(u >> 32));
    return u2d(((uint64_t)hi << 32) | lo);
}
template <int L, bool UP, int OFF = 1>
__device__ __forceinline__ double xlane_sum(double s) {
    if constexpr (OFF < L) {
        const double o = xchg_d<OFF>(s);
        return xlane_sum<L, UP, 2 * OFF>(UP ? add_ru(s, o) : add_rd(s, o));
    } else {
        return s;
    }
}
template <int N, int L, bool UP, bool U>
__device__ __forceinline__ double psum_arr(const double *v) {
    constexpr int R = N / L;
    double s;
    if constexpr (R == 1) {
        s = v[0];
    } else if constexpr (U) {
        double t[R];
#pragma unroll
        for (int r = 0; r < R; r++) t[r] = v[r];
#pragma unroll
        for (int w = 1; w < R; w <<= 1)
#pragma unroll
            for (int k = 0; k < R; k += 2 * w) t[k] = UP ? add_ru(t[k], t[k + w]) : add_rd(t[k], t[k + w]);
        s = t[0];
    } else {
        double t[R];
        for (int r = 0; r < R; r++) t[r] = v[r];
        for (int w = 1; w < R; w <<= 1)
            for (int k = 0; k < R; k += 2 * w) t[k] = UP ? add_ru(t[k], t[k + w]) : add_rd(t[k], t[k + w]);
        s = t[0];
    }
    return xlane_sum<L, UP>(s);
}
template <int L, int OFF = 1>
__device__ __forceinline__ uint32_t xlane_usum(uint32_t v) {
    if constexpr (OFF < L)
        return xlane_usum<L, 2 * OFF>(v + (uint32_t)xchg<OFF>((int)v));
    else
        return v;
}
template <int L, int OFF = 1>
__device__ __forceinline__ int xlane_imax(int v) {
    if constexpr (OFF < L) {
        const int o = xchg<OFF>(v);
        return xlane_imax<L, 2 * OFF>(o > v ? o : v);
    } else {
        return v;
    }
}
template <int L>
__device__ __forceinline__ int bcast(int v, int src) {
    if constexpr (L == 1)
        return v;
    else
        return __shfl(v, src, L);
}

// ---------------- mixed-radix conversion ----------------
template <int N, int L, bool U>
__device__ __forceinline__ void rns_mrc(Core<N, L> k, int *mr, const int *x) {
    constexpr int R = N / L;
    loop<R, U>([&](int r) { mr[r] = x[r]; });
    loop<N - 1, U>([&](int j) {
        int own;
        if constexpr (U) {
            own = mr[j % R];
        } else {
            own = (k.lane == j / R) ? mr[j % R] : 0;
        }
        const int v = bcast<L>(own, j / R);
        loop<R, U>([&](int r) {
            const int i = k.ix(r);
            if (i > j) {
                int t = mr[r] - v;
                t = (mr[r] < v) ? t + k.mi(r) : t;
                mr[r] = mulmod(t, k.mrc(j, r), k.md(r), k.mv(r));
            }
        });
    });
}
template <int N, int L, bool U>
__device__ __forceinline__ int rns_mrd(Core<N, L> k, const int *x) {
    constexpr int R = N / L;
    int mr[R];
    rns_mrc<N, L, U>(k, mr, x);
    return bcast<L>(mr[R - 1], L - 1);
}
template <int N, int L, bool U>
__device__ __forceinline__ int rns_mrc_compare(Core<N, L> k, const int *x, const int *y) {
    constexpr int R = N / L;
    int mx[R], my[R];
    rns_mrc<N, L, U>(k, mx, x);
    rns_mrc<N, L, U>(k, my, y);
    int key = -1;
    loop<R, U>([&](int r) {
        if (mx[r] != my[r]) key = 2 * k.ix(r) + (mx[r] > my[r] ? 1 : 0);
    });
    key = xlane_imax<L>(key);
    return key < 0 ? 0 : ((key & 1) ? 1 : -1);
}

// ---------------- interval evaluation ----------------
constexpr int EVAL_MAX_ITER = 128;  // bound on the refinement loop; valid inputs need a handful

template <int N, int L, bool U, bool FULL>
__device__ __forceinline__ void rns_eval(Core<N, L> k, const int *x, Er &lo, Er &up, bool known_small) {
    constexpr int R = N / L;
    const DevConsts *C = k.C;
    int s[R];
    double fl[R], fu[R];
    loop<R, U>([&](int r) {
        s[r] = mulmod(x[r], k.w(r), k.md(r), k.mv(r));
        fl[r] = mul_rd((double)s[r], k.rrd(r));
        fu[r] = mul_ru((double)s[r], k.rru(r));
    });
    double suml = psum_arr<N, L, false, U>(fl);
    double sumu = psum_arr<N, L, true, U>(fu);
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
        if (whl != whu) mr = known_small ? 0 : rns_mrd<N, L, U>(k, x);
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
        int kk = -(ceil_log2_rn(sumu) + 1);
        kk = kk > ref ? kk : ref;
        kk = kk < kmax ? kk : kmax;
        const int *row = k.P2 + (size_t)kk * N;
        loop<R, U>([&](int r) {
            s[r] = mulmod(s[r], row[k.ix(r)], k.md(r), k.mv(r));
            fu[r] = mul_ru((double)s[r], k.rru(r));
        });
        sumu = psum_arr<N, L, true, U>(fu);
        sumu = sub_ru(sumu, (double)(unsigned)sumu);
        K += kk;
    }
    loop<R, U>([&](int r) { fl[r] = mul_rd((double)s[r], k.rrd(r)); });
    suml = psum_arr<N, L, false, U>(fl);
    suml = sub_rd(suml, (double)(unsigned)suml);
    lo = er_set_d(suml);
    up = er_set_d(sumu);
    lo.exp -= K;
    up.exp -= K;
}

// ---------------- power-of-two scaling ----------------
template <int N, int L, bool U>
__device__ __forceinline__ int rns_rank_full(Core<N, L> k, const int *x, const int *c) {
    constexpr int R = N / L;
    double fl[R], fu[R];
    loop<R, U>([&](int r) {
        fl[r] = mul_rd((double)c[r], k.rrd(r));
        fu[r] = mul_ru((double)c[r], k.rru(r));
    });
    const unsigned whl = (unsigned)psum_arr<N, L, false, U>(fl), whu = (unsigned)psum_arr<N, L, true, U>(fu);
    if (whl == whu) return (int)whl;
    return rns_mrd<N, L, U>(k, x) == 0 ? (int)whu : (int)whl;
}
template <int N, int L, bool U>
__device__ __forceinline__ int rns_rank_fast(Core<N, L> k, const int *c) {
    constexpr int R = N / L;
    double fu[R];
    loop<R, U>([&](int r) { fu[r] = mul_ru((double)c[r], k.rru(r)); });
    return (int)psum_arr<N, L, true, U>(fu);
}
// y = (x - (X mod 2^j)) * 2^-j, given the rank kr of x
template <int N, int L, bool U>
__device__ __forceinline__ void rns_scale_step(Core<N, L> k, int *y, int kr, int j, const int *x, const int *c) {
    constexpr int R = N / L;
    const uint32_t mask = (1u << j) - 1u;
    uint32_t part = 0;
    loop<R, U>([&](int r) { part += ((uint32_t)k.C->h[j - 1][k.ix(r)] * (uint32_t)c[r]) & mask; });
    uint32_t res = xlane_usum<L>(part);
    res = (res - (uint32_t)kr * (uint32_t)k.C->mpow2[j - 1]) & mask;
    loop<R, U>([&](int r) {
        const int mult = modd((double)res, k.md(r), k.mv(r));
        int t = x[r] - mult;
        t = t < 0 ? t + k.mi(r) : t;
        y[r] = mulmod(t, k.C->pow2inv[j - 1][k.ix(r)], k.md(r), k.mv(r));
    });
}
template <int N, int L, bool U>
__device__ __forceinline__ void rns_xw(Core<N, L> k, int *c, const int *x) {
    constexpr int R = N / L;
    loop<R, U>([&](int r) { c[r] = mulmod(x[r], k.w(r), k.md(r), k.mv(r)); });
}
template <int N, int L, bool U>
__device__ __forceinline__ void rns_scale2pow(Core<N, L> k, int *x, unsigned D) {
    constexpr int R = N / L;
    int c[R];
    int t = (int)(D / P2T);
    if (t > 0) {
        rns_xw<N, L, U>(k, c, x);
        const int kr = rns_rank_full<N, L, U>(k, x, c);
        rns_scale_step<N, L, U>(k, x, kr, P2T, x, c);
        t -= 1;
    }
    while (t > 0) {
        rns_xw<N, L, U>(k, c, x);
        const int kr = rns_rank_fast<N, L, U>(k, c);
        rns_scale_step<N, L, U>(k, x, kr, P2T, x, c);
        t -= 1;
    }
    const unsigned d = D % P2T;
    if (d > 0) {
        rns_xw<N, L, U>(k, c, x);
        const int kr = (d < D) ? rns_rank_full<N, L, U>(k, x, c) : rns_rank_fast<N, L, U>(k, c);
        rns_scale_step<N, L, U>(k, x, kr, (int)d, x, c);
    }
}

// mp_round body: exp += bits, X = floor(X / 2^bits), fast re-evaluation
template <int N, int L, bool U>
__device__ __forceinline__ void mp_round_body(Core<N, L> k, int *d, int &exp, Er &ev0, Er &ev1) {
    const int bits = ev1.exp - k.C->mp_h + 1;
    exp += bits;
    rns_scale2pow<N, L, U>(k, d, (unsigned)bits);
    rns_eval<N, L, U, false>(k, d, ev0, ev1, false);
}

// ---------------- out-of-line entry points (U = false) ----------------
template <int N, int L>
__device__ __noinline__ void ool_round(Core<N, L> k, int *d, int *exp, Er *ev0, Er *ev1) {
    int e = *exp;
    Er a = *ev0, b = *ev1;
    mp_round_body<N, L, false>(k, d, e, a, b);
    *exp = e;
    *ev0 = a;
    *ev1 = b;
}
template <int N, int L>
__device__ __noinline__ int ool_mrc_compare(Core<N, L> k, const int *x, const int *y) {
    return rns_mrc_compare<N, L, false>(k, x, y);
}
template <int N, int L>
__device__ __noinline__ void ool_eval_full(Core<N, L> k, const int *x, Er *lo, Er *up, bool known_small) {
    Er a, b;
    rns_eval<N, L, false, true>(k, x, a, b, known_small);
    *lo = a;
    *up = b;
}

}  // namespace mph
