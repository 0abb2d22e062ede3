// Kernel templates of the multiple-precision path, parameterised by the
// number of moduli N and the lanes per number L (see group.hpp).
//
// Thread mapping: blocks of BLK threads; every L consecutive lanes form one
// group, i.e. one multiple-precision number. The reference's "one thread per
// element" launch shapes are replaced by this mapping; the reference's launch
// parameters survive only where they fix a summation order (dot: gridDim3 x
// blockDim3, gemv: blockDim3, SpMV CSR-vector: threadsPerRow), and there the
// same order is reproduced.
//
// Reductions over 16 and 32 moduli run in two kernels. The fast kernel
// (FULL = false) carries only the common case of a multiply-accumulate step:
// no rounding and an interval evaluation of the sum that does not straddle
// zero. The rare cases need mixed-radix conversion, whose registers would
// otherwise be reserved for the whole loop. A chain that meets one is
// abandoned and marked (sign = REDO in its workspace result); the fixup
// kernel (FULL = true, FIX = true) recomputes exactly the marked chains with
// every case of the reference arithmetic. Both kernels run the identical
// sequence of operations, so the result does not depend on which one
// finished a chain. Over 8 moduli rounding is the common case and the
// single full kernel is used.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "group.hpp"
#include "hostops.hpp"
#include "ops.hpp"

namespace mph {

constexpr int BLK = 256;
constexpr int REDO = 2;  // sign of a workspace result whose chain the fixup kernel recomputes


template <int N, int L>
__device__ __forceinline__ void load_soa(const Grp<N, L> &g, MP<N / L> &a, const SoA &v, long long i) {
    constexpr int R = N / L;
    const int *p = v.d + i * N + g.lane() * R;
#pragma unroll
    for (int r = 0; r < R; r++) a.d[r] = p[r];
    a.sign = v.s[i];
    a.exp = v.e[i];
    const ErMem lo = v.ev[i], up = v.ev[i + v.len];
    a.ev0 = Er{lo.frac, (int)lo.exp};
    a.ev1 = Er{up.frac, (int)up.exp};
}
template <int N, int L>
__device__ __forceinline__ void store_soa(const Grp<N, L> &g, const MP<N / L> &a, const SoA &v, long long i) {
    constexpr int R = N / L;
    int *p = v.d + i * N + g.lane() * R;
#pragma unroll
    for (int r = 0; r < R; r++) p[r] = a.d[r];
    if (g.lane() == 0) {
        v.s[i] = a.sign;
        v.e[i] = a.exp;
        v.ev[i] = ErMem{a.ev0.frac, (long long)a.ev0.exp};
        v.ev[i + v.len] = ErMem{a.ev1.frac, (long long)a.ev1.exp};
    }
}
template <int N, int L>
__device__ __forceinline__ void load_aos(const Grp<N, L> &g, MP<N / L> &a, const AoSRef<N> &v, long long i) {
    constexpr int R = N / L;
    const int *p = v.digits(i) + g.lane() * R;
#pragma unroll
    for (int r = 0; r < R; r++) a.d[r] = p[r];
    a.sign = *v.sign(i);
    a.exp = *v.exp(i);
    const ErMem *e = v.eval(i);
    const ErMem lo = e[0], up = e[1];
    a.ev0 = Er{lo.frac, (int)lo.exp};
    a.ev1 = Er{up.frac, (int)up.exp};
}
template <int N, int L>
__device__ __forceinline__ void store_aos(const Grp<N, L> &g, const MP<N / L> &a, const AoSRef<N> &v, long long i) {
    constexpr int R = N / L;
    int *p = v.digits(i) + g.lane() * R;
#pragma unroll
    for (int r = 0; r < R; r++) p[r] = a.d[r];
    if (g.lane() == 0) {
        *v.sign(i) = a.sign;
        *v.exp(i) = a.exp;
        ErMem *e = v.eval(i);
        e[0] = ErMem{a.ev0.frac, (long long)a.ev0.exp};
        e[1] = ErMem{a.ev1.frac, (long long)a.ev1.exp};
    }
}

// BLAS index of element k of an n-vector with increment inc (e.g. src/mpvector.cuh:65-67)
__device__ __forceinline__ long long bidx(long long k, long long n, int inc) {
    return inc > 0 ? k * inc : (-n + k + 1) * (long long)inc;
}

// read the device-side length of an mp_array_t argument
__device__ __forceinline__ SoA rd(SoA v) {
    v.len = (v.lenp != nullptr) ? *v.lenp : v.len;
    return v;
}

template <int L>
__device__ __forceinline__ long long group_id() {
    return ((long long)blockIdx.x * BLK + threadIdx.x) / L;
}

// ---------------------------------------------------------------- conversion
// mp_set_d for count doubles; out either SoA (soa_out) or AoS
template <int N, int L, bool SOA_OUT>
__global__ __launch_bounds__(BLK) void k_set_d(const DevConsts *__restrict__ C, const double *__restrict__ x, int count,
                                               SoA so, AoSRef<N> ao) {
    so = rd(so);
    const long long i = group_id<L>();
    if (i >= count) return;
    Grp<N, L> g(C);
    MP<N / L> a;
    g.set_d(a, x[i]);
    if (SOA_OUT)
        store_soa(g, a, so, i);
    else
        store_aos(g, a, ao, i);
}

// sign and exponent of element i in either layout
template <int N>
__device__ __forceinline__ int sign_at(const SoA &v, long long i) { return v.s[i]; }
template <int N>
__device__ __forceinline__ int sign_at(const AoSRef<N> &v, long long i) { return *v.sign(i); }
template <int N>
__device__ __forceinline__ int exp_of(const SoA &v, long long i) { return v.e[i]; }
template <int N>
__device__ __forceinline__ int exp_of(const AoSRef<N> &v, long long i) { return *v.exp(i); }

// ---------------------------------------------------------------- export
// The integer X (0 <= X < M) that a number's residues stand for, as the reference's
// mp_get_d / mp_get_mpfr form it (src/arith/assign.cuh:154-211, CRT with the orthogonal
// base, mod M): here by mixed-radix conversion, X = a_0 + a_1 m_0 + a_2 m_0 m_1 + ...,
// then Horner's rule over 32-bit limbs, least significant first. Residues outside
// [0, m_i) are reduced first. One thread per number, loops unrolled (arrays in registers).
template <int N>
constexpr int export_limbs() {
    return (31 * N + 31) / 32 + 1;  // moduli below 2^31: X < 2^(31 N)
}
template <int N>
__device__ __forceinline__ void mp_integer(const DevConsts *__restrict__ C, const int *d, uint32_t *X) {
    constexpr int LW = export_limbs<N>();
    int a[N];
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int m = C->m[i];
        int v = d[i] % m;
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
#pragma unroll
    for (int l = 0; l < LW; l++) X[l] = 0;
    X[0] = (uint32_t)a[N - 1];
#pragma unroll
    for (int i = N - 2; i >= 0; i--) {
        uint64_t carry = (uint64_t)a[i];
        const uint64_t m = (uint64_t)C->m[i];
#pragma unroll
        for (int l = 0; l < LW; l++) {
            const uint64_t t = (uint64_t)X[l] * m + carry;
            X[l] = (uint32_t)t;
            carry = t >> 32;
        }
    }
}
// (-1)^sign X 2^exp as a double: X rounded to 53 significant bits to nearest (ties to
// even), then scaled by a power of two with one more rounding to nearest when the result
// is subnormal (overflow gives an infinity) - the reference's mpfr_set_str at MPFR's
// default 53-bit precision followed by mpfr_get_d (src/arith/assign.cuh:167-175).
template <int N>
__device__ __forceinline__ double integer_to_double(const uint32_t *X, int sign, int exp) {
    constexpr int LW = export_limbs<N>();
    int top = -1;
#pragma unroll
    for (int l = 0; l < LW; l++) top = X[l] ? l : top;
    if (top < 0) return sign ? -0.0 : 0.0;
    uint32_t t0 = 0, t1 = 0, t2 = 0;  // limbs top, top - 1, top - 2
    bool below = false;               // any nonzero limb under top - 2
#pragma unroll
    for (int l = 0; l < LW; l++) {
        t0 = (l == top) ? X[l] : t0;
        t1 = (l == top - 1) ? X[l] : t1;
        t2 = (l == top - 2) ? X[l] : t2;
        below = below || (l < top - 2 && X[l] != 0);
    }
    const int lz = __clz(t0);
    uint64_t w = ((uint64_t)t0 << 32) | t1;  // the 64 bits from X's leading one down
    bool sticky = below;
    if (lz) {
        w = (w << lz) | (t2 >> (32 - lz));
        sticky = sticky || (t2 << lz) != 0;
    } else {
        sticky = sticky || t2 != 0;
    }
    uint64_t sig = w >> 11;
    const uint32_t rem = (uint32_t)(w & 0x7ff);
    sticky = sticky || (rem & 0x3ff) != 0;
    if ((rem & 0x400) && (sticky || (sig & 1))) sig++;  // may reach 2^53, still exact
    const int e2 = 32 * top + (32 - lz) - 53 + exp;
    const double r = ldexp((double)sig, e2);
    return sign ? -r : r;
}
template <int N>
__device__ __forceinline__ void gather_digits(const SoA &v, long long i, int *d) {
#pragma unroll
    for (int r = 0; r < N; r++) d[r] = v.d[i * N + r];
}
template <int N>
__device__ __forceinline__ void gather_digits(const AoSRef<N> &v, long long i, int *d) {
#pragma unroll
    for (int r = 0; r < N; r++) d[r] = v.digits(i)[r];
}
// mp_get_d over count numbers (MODE 0), or the exact integers (MODE 1): export_limbs<N>()
// limbs per number, with the sign and the exponent
template <int N, int L, class VM, int MODE>
__global__ __launch_bounds__(BLK) void k_export(const DevConsts *__restrict__ C, int count, VM x,
                                               double *__restrict__ out, uint32_t *__restrict__ limbs,
                                               int *__restrict__ sg, int *__restrict__ ex) {
    if constexpr (std::is_same<VM, SoA>::value) x = rd(x);
    const long long i = (long long)blockIdx.x * BLK + threadIdx.x;
    if (i >= count) return;
    constexpr int LW = export_limbs<N>();
    int d[N];
    gather_digits<N>(x, i, d);
    uint32_t X[LW];
    mp_integer<N>(C, d, X);
    const int s = sign_at<N>(x, i), e = exp_of<N>(x, i);
    if constexpr (MODE == 0) {
        out[i] = integer_to_double<N>(X, s, e);
    } else {
#pragma unroll
        for (int l = 0; l < LW; l++) limbs[i * LW + l] = X[l];
        sg[i] = s;
        ex[i] = e;
    }
}

// ---------------------------------------------------------------- host MPFR operations
// mp_div (src/arith/div.cuh:31-65) on count numbers: r[i] = x[i] / y[i] as MPFR computes it
// at MP_PRECISION bits, with the CPU evaluation (hostops.hpp). A zero divisor gives zero
// (the reference's result is undefined).
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_div(const DevConsts *__restrict__ C, int count, AoSRef<N> x, AoSRef<N> y,
                                            AoSRef<N> r) {
    const long long i = (long long)blockIdx.x * BLK + threadIdx.x;
    if (i >= count) return;
    int dx[N], dy[N];
    gather_digits<N>(x, i, dx);
    gather_digits<N>(y, i, dy);
    const int sign = *x.sign(i) ^ *y.sign(i);
    int ea, eb, e = 0;
    const BigN<N> a = mp_to_big<N>(C, dx, *x.exp(i), ea), b = mp_to_big<N>(C, dy, *y.exp(i), eb);
    BigN<N> q;
    big_zero(q);
    if (!big_is_zero(a) && !big_is_zero(b)) q = big_quotient<N>(a, ea, b, eb, C->mp_precision, e);
    int dr[N], sr, er;
    Er lo, up;
    big_to_mp<N>(C, q, e, sign, dr, sr, er, lo, up);
#pragma unroll
    for (int k = 0; k < N; k++) r.digits(i)[k] = dr[k];
    *r.sign(i) = sr;
    *r.exp(i) = er;
    r.eval(i)[0] = ErMem{lo.frac, (long long)lo.exp};
    r.eval(i)[1] = ErMem{up.frac, (long long)up.exp};
}
// mp_sqrt (MODE 0: multiple-precision result) and mp_dsqrt (MODE 1: double), src/arith/
// sqrt.cuh:33-89. A negative operand gives NaN (MODE 1) or zero (MODE 0); MPFR gives NaN.
template <int N, int L, int MODE>
__global__ __launch_bounds__(BLK) void k_sqrt(const DevConsts *__restrict__ C, int count, AoSRef<N> x, AoSRef<N> r,
                                             double *__restrict__ out) {
    const long long i = (long long)blockIdx.x * BLK + threadIdx.x;
    if (i >= count) return;
    int dx[N];
    gather_digits<N>(x, i, dx);
    const int sign = *x.sign(i);
    int ea, e = 0;
    const BigN<N> a = mp_to_big<N>(C, dx, *x.exp(i), ea);
    const bool zero = big_is_zero(a);
    Big<2 * big_limbs<N>()> s;
    big_zero(s);
    if (!zero && !sign) s = big_sqrt_round<N>(a, ea, C->mp_precision, e);
    if constexpr (MODE == 1) {
        out[i] = (sign && !zero) ? __builtin_nan("") : big_to_double(s, e, sign);
    } else {
        BigN<N> q;
#pragma unroll
        for (int k = 0; k < big_limbs<N>(); k++) q.l[k] = s.l[k];  // the root has MP_PRECISION bits
        int dr[N], sr, er;
        Er lo, up;
        big_to_mp<N>(C, q, e, sign && zero, dr, sr, er, lo, up);
#pragma unroll
        for (int k = 0; k < N; k++) r.digits(i)[k] = dr[k];
        *r.sign(i) = sr;
        *r.exp(i) = er;
        r.eval(i)[0] = ErMem{lo.frac, (long long)lo.exp};
        r.eval(i)[1] = ErMem{up.frac, (long long)up.exp};
    }
}

// ---------------------------------------------------------------- element-wise
template <int N, int L, int OP>
__global__ __launch_bounds__(BLK) void k_elem(const DevConsts *__restrict__ C, int count, AoSRef<N> x, AoSRef<N> y,
                                              const double *__restrict__ yd, AoSRef<N> r) {
    const long long i = group_id<L>();
    if (i >= count) return;
    Grp<N, L> g(C);
    MP<N / L> a, b, c;
    load_aos(g, a, x, i);
    if (OP == 3) {
        g.mul_d(c, a, yd[i]);
    } else {
        load_aos(g, b, y, i);
        if (OP == 0) g.add(c, a, b);
        if (OP == 1) {
            b.sign ^= 1;  // mp_sub, src/arith/sub.cuh:41-53
            g.add(c, a, b);
        }
        if (OP == 2) g.mul(c, a, b);
        if (OP == 4) g.vadd(c, a, b);
    }
    store_aos(g, c, r, i);
}

// comparisons: r[i] in {-1, 0, 1}. MODE 0: cuda::mp_cmp (src/arith/cmp.cuh:120-124), 1:
// cuda::mp_cmp_abs (src/arith/cmpabs.cuh:115-119), both on AoS numbers; 2: cuda::rns_cmp on
// bare residues, count x N ints each (src/rns.cuh:1242-1260)
template <int N, int L, int MODE>
__global__ __launch_bounds__(BLK) void k_cmp(const DevConsts *__restrict__ C, int count, AoSRef<N> x, AoSRef<N> y,
                                             const int *__restrict__ dx, const int *__restrict__ dy,
                                             int *__restrict__ out) {
    const long long i = group_id<L>();
    if (i >= count) return;
    Grp<N, L> g(C);
    int c;
    if constexpr (MODE == 2) {
        constexpr int R = N / L;
        int a[R], b[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            a[r] = dx[i * N + g.ix(r)];
            b[r] = dy[i * N + g.ix(r)];
        }
        c = g.rns_cmp(a, b);
    } else {
        MP<N / L> a, b;
        load_aos(g, a, x, i);
        load_aos(g, b, y, i);
        c = g.template cmp<MODE == 1>(a, b);
    }
    if (g.lane() == 0) out[i] = c;
}

// ---------------------------------------------------------------- reductions
// Strided fold of the reduction kernels (src/mpreduct.cuh:38-56, dot_v2.cuh:37-55):
// virtual thread v accumulates terms v, v + nvt, v + 2 nvt, ... from zero.
// MODE 0: terms are x_i * y_i with SoA x, y and BLAS strides (v1 dot)
// MODE 1: terms are x_i * y_i with AoS x, y (v2 dot)
// MODE 2: terms are AoS partials t_i (always FULL, never FIX)
// MODE 3: terms are |x_i| with SoA x and increment incx > 0 (mp_array_reduce_sum_abs_kernel1,
//         src/mpreduct.cuh:120-153: mp_add_abs of nonnegative operands is mp_add; always FULL)
template <int N, int L, int MODE, bool FULL, bool FIX>
__global__ __launch_bounds__(BLK) void k_fold(const DevConsts *__restrict__ C, long long cnt, long long nvt, SoA xs,
                                              int incx, SoA ys, int incy, AoSRef<N> xa, AoSRef<N> ya, AoSRef<N> out) {
    xs = rd(xs);
    ys = rd(ys);
    const long long v = group_id<L>();
    if (v >= nvt) return;
    if (FIX && *out.sign(v) != REDO) return;
    constexpr int R = N / L;
    Grp<N, L> g(C);
    MP<R> acc, p;
    g.zero(acc);
    bool ok = true;
    if constexpr (MODE == 0 || MODE == 1) {
        // products x_i * y_i: the operands of term i + nvt are loaded while step i runs, so
        // the chain's steps do not wait on memory (Grp::fma_step's arithmetic, unchanged)
        struct Op {
            int xs, xe, ys, ye;
            Er x0, x1, y0, y1;
            int dx[R], dy[R];
        };
        auto load = [&](long long i) {
            Op o;
            const int *px, *py;
            if constexpr (MODE == 0) {
                const long long ix = bidx(i, cnt, incx), iy = bidx(i, cnt, incy);
                scalars_of<N>(xs, ix, o.xs, o.xe, o.x0, o.x1);
                scalars_of<N>(ys, iy, o.ys, o.ye, o.y0, o.y1);
                px = digits_of<N>(xs, ix);
                py = digits_of<N>(ys, iy);
            } else {
                scalars_of<N>(xa, i, o.xs, o.xe, o.x0, o.x1);
                scalars_of<N>(ya, i, o.ys, o.ye, o.y0, o.y1);
                px = digits_of<N>(xa, i);
                py = digits_of<N>(ya, i);
            }
#pragma unroll
            for (int r = 0; r < R; r++) {
                o.dx[r] = px[g.ix(r)];
                o.dy[r] = py[g.ix(r)];
            }
            return o;
        };
        if (v < cnt) {
            Op cur = load(v);
            for (long long i = v; i < cnt; i += nvt) {
                const Op nxt = load(i + nvt < cnt ? i + nvt : i);
                g.mul_scalars(p, cur.xs, cur.xe, cur.x0, cur.x1, cur.ys, cur.ye, cur.y0, cur.y1);
                ok = g.template step_tail<FULL>(acc, p, cur.dx, cur.dy, 0.0, false);
                if (!ok) break;
                cur = nxt;
            }
        }
    } else {
        for (long long i = v; i < cnt; i += nvt) {
            if (MODE == 3) {
                load_soa(g, p, xs, i * incx);
                p.sign = 0;
            } else {
                load_aos(g, p, xa, i);
            }
            g.add(acc, acc, p);
        }
    }
    if (ok)
        store_aos(g, acc, out, v);
    else if (g.lane() == 0)
        *out.sign(v) = REDO;
}

// Halving tree of src/mpreduct.cuh:60-68 over each block of b partials:
// for i = nextPow2(b)/2 .. 1: s[t] = s[t] + s[t+i] for t < i, t + i < b.
// One workgroup per block; the block's partials stay in (cached) global memory.
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_tree(const DevConsts *__restrict__ C, AoSRef<N> part, int b, unsigned p2,
                                              AoSRef<N> out) {
    const long long base = (long long)blockIdx.x * b;
    const int W = BLK / L;
    const int gl = threadIdx.x / L;
    Grp<N, L> g(C);
    for (unsigned s = p2 >> 1; s >= 1; s >>= 1) {
        for (int t = gl; t < (int)s; t += W) {
            if (t + (int)s < b) {
                MP<N / L> a, c, r;
                load_aos(g, a, part, base + t);
                load_aos(g, c, part, base + t + s);
                g.add(r, a, c);
                store_aos(g, r, part, base + t);
            }
        }
        __threadfence_block();
        __syncthreads();
    }
    if (gl == 0) {
        MP<N / L> a;
        load_aos(g, a, part, base);
        store_aos(g, a, out, blockIdx.x);
    }
}

// AoS -> SoA copy of count numbers (also used to write reduction results into mp_array_t)
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_aos_to_soa(const DevConsts *__restrict__ C, AoSRef<N> src, SoA dst, int count) {
    dst = rd(dst);
    const long long i = group_id<L>();
    if (i >= count) return;
    Grp<N, L> g(C);
    MP<N / L> a;
    load_aos(g, a, src, i);
    store_soa(g, a, dst, i);
}
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_soa_to_aos(const DevConsts *__restrict__ C, SoA src, AoSRef<N> dst, int count) {
    src = rd(src);
    const long long i = group_id<L>();
    if (i >= count) return;
    Grp<N, L> g(C);
    MP<N / L> a;
    load_soa(g, a, src, i);
    store_aos(g, a, dst, i);
}

// ---------------------------------------------------------------- axpy
// v1: buffer = x * alpha (+round), y = buffer (+) y with the v1 vector add (+round),
// src/blas/axpy.cuh:62-72 fused into one pass
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_axpy(const DevConsts *__restrict__ C, int n, SoA alpha, SoA x, int incx, SoA y,
                                              int incy) {
    alpha = rd(alpha);
    x = rd(x);
    y = rd(y);
    const long long i = group_id<L>();
    if (i >= n) return;
    Grp<N, L> g(C);
    MP<N / L> a, xv, yv, b, r;
    load_soa(g, a, alpha, 0);
    load_soa(g, xv, x, bidx(i, n, incx));
    const long long iy = bidx(i, n, incy);
    load_soa(g, yv, y, iy);
    g.mul(b, xv, a);
    g.vadd(r, b, yv);
    store_soa(g, r, y, iy);
}
// v2: r = alpha * x (+) y, src/blas/v2/axpy_v2.cuh:39-52
// NEG: mp_maxpy, r = (-alpha) * x (+) y (src/blas/v2/maxpy_v2.cuh)
template <int N, int L, bool NEG = false>
__global__ __launch_bounds__(BLK) void k_axpy_v2(const DevConsts *__restrict__ C, int n, AoSRef<N> alpha, AoSRef<N> x,
                                                 AoSRef<N> y, AoSRef<N> r) {
    const long long i = group_id<L>();
    if (i >= n) return;
    Grp<N, L> g(C);
    MP<N / L> a, xv, yv, ax, s;
    load_aos(g, a, alpha, 0);
    if (NEG) a.sign ^= 1;
    load_aos(g, xv, x, i);
    load_aos(g, yv, y, i);
    g.mul(ax, a, xv);
    g.add(s, ax, yv);
    store_aos(g, s, r, i);
}

// ---------------------------------------------------------------- gemv
// out[k] = v[k] * s (vec2scal esi/digits/round, src/mpvector.cuh:139-216), v with BLAS stride;
// out is AoS (ws) or the SoA v itself (in place)
template <int N, int L, bool INPLACE>
__global__ __launch_bounds__(BLK) void k_scal(const DevConsts *__restrict__ C, int len, SoA s, SoA v, int inc,
                                              AoSRef<N> out) {
    s = rd(s);
    v = rd(v);
    const long long k = group_id<L>();
    if (k >= len) return;
    Grp<N, L> g(C);
    MP<N / L> a, x, r;
    load_soa(g, a, s, 0);
    const long long iv = bidx(k, len, inc);
    load_soa(g, x, v, iv);
    g.mul(r, x, a);
    if (INPLACE)
        store_soa(g, r, v, iv);
    else
        store_aos(g, r, out, k);
}
// virtual thread (o, t): o-th output, t-th of b3 strided partial sums of A-entries times bx
// (matrix_row_sum_kernel / matrix_col_sum_kernel, src/blas/gemv.cuh:44-120, with the
// products of the mat2vec scaling kernels formed on the fly); its partial goes to
// part[o * b3 + t]. Groups are laid out so that a wavefront's loads are contiguous:
// without transpose consecutive groups take consecutive rows o (same t), so each
// step reads a run of one column of A; with transpose consecutive t of one output.
template <int N, int L, bool TRANS, bool FULL, bool FIX>
__global__ __launch_bounds__(BLK) void k_gemv_fold(const DevConsts *__restrict__ C, int m, int n, SoA A, int lda,
                                                   AoSRef<N> bx, int b3, long long nvt, AoSRef<N> part) {
    A = rd(A);
    const long long gid = group_id<L>();
    if (gid >= nvt) return;
    const long long ly = nvt / b3;
    const long long o = TRANS ? gid / b3 : gid % ly;
    const int t = (int)(TRANS ? gid % b3 : gid / ly);
    const long long v = o * b3 + t;
    if (FIX && *part.sign(v) != REDO) return;
    Grp<N, L> g(C);
    MP<N / L> acc;
    g.zero(acc);
    const int len = TRANS ? m : n;
    bool ok = true;
    for (int j = t; j < len && ok; j += b3)
        ok = g.template fma_step<FULL>(acc, A, TRANS ? (o * lda + j) : ((long long)j * lda + o), bx, j);
    if (ok)
        store_aos(g, acc, part, v);
    else if (g.lane() == 0)
        *part.sign(v) = REDO;
}
// Residues of a run of `rows` consecutive elements of an SoA array (rows * N ints,
// 16-byte aligned) moved to LDS with element stride N + 1. When a full run
// (rows == W) is a whole number of 16-byte loads per thread (VEC), it goes
// through registers in two halves, so that the next step's loads can be issued
// before the current step computes: fetch issues the loads, store writes them
// to LDS. Other runs are copied directly by store.
// T: threads of the workgroup (BLK unless the kernel launches fewer).
template <int N, int L, int T = BLK>
struct RunStage {
    static constexpr int W = T / L, SP = N + 1;
    static constexpr bool VEC = (W * N) % (4 * T) == 0;
    static constexpr int V4 = VEC ? W * N / 4 / T : 1;  // int4 per thread for a full run
    int4 v[V4];
    __device__ __forceinline__ void fetch(const int *src, int rows) {
        if (VEC && rows == W) {
            const int4 *s4 = (const int4 *)src;
#pragma unroll
            for (int q = 0; q < V4; q++) v[q] = s4[q * T + threadIdx.x];
        }
    }
    __device__ __forceinline__ void store(int *sd, const int *src, int rows) const {
        if (VEC && rows == W) {
#pragma unroll
            for (int q = 0; q < V4; q++) {
                const int i = 4 * (q * T + threadIdx.x), e = i / N, f = i % N;
                int *d = sd + e * SP + f;
                d[0] = v[q].x;
                d[1] = v[q].y;
                d[2] = v[q].z;
                d[3] = v[q].w;
            }
        } else {
            for (int q = threadIdx.x; q < rows * N; q += T) sd[(q / N) * SP + q % N] = src[q];
        }
    }
};

// Main pass without transpose, the residues of A staged through LDS: workgroup
// (x, t) holds the W = BLK / L consecutive rows o of block x, all with strided
// index t, so each step j reads one contiguous run of column j of A. The run's
// residues are loaded cooperatively (coalesced) into LDS with a padded stride and
// every group reads its own from there. Same partials as k_gemv_fold.
template <int N, int L, bool FULL>
__global__ __launch_bounds__(BLK) void k_gemv_rows(const DevConsts *__restrict__ C, int m, int n, SoA A, int lda,
                                                   AoSRef<N> bx, int b3, AoSRef<N> part) {
    constexpr int W = BLK / L, R = N / L, SP = N + 1;
    __shared__ int sd[W * SP];
    A = rd(A);
    const int gl = threadIdx.x / L;
    const int t = blockIdx.y;
    const long long o0 = (long long)blockIdx.x * W, o = o0 + gl;
    const int rows = (int)((m - o0) < W ? (m - o0) : W);
    bool ok = o < m;
    Grp<N, L> g(C);
    MP<R> acc;
    g.zero(acc);
    RunStage<N, L> stg;
    if (t < n) stg.fetch(A.d + (o0 + (long long)t * lda) * N, rows);
    for (int j = t; j < n; j += b3) {
        stg.store(sd, A.d + (o0 + (long long)j * lda) * N, rows);
        __syncthreads();
        if (j + b3 < n) stg.fetch(A.d + (o0 + (long long)(j + b3) * lda) * N, rows);
        if (ok) ok = g.template fma_step_d<FULL>(acc, A, o + (long long)j * lda, sd + gl * SP + g.lane() * R, bx, j);
        __syncthreads();
    }
    if (o >= m) return;
    const long long v = o * b3 + t;
    if (ok)
        store_aos(g, acc, part, v);
    else if (g.lane() == 0)
        *part.sign(v) = REDO;
}

// y[o] = y[o] (+) rowsum[o]  (src/blas/gemv.cuh:72-75 with mp_add rounding, add.cuh:223-248)
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_gemv_final(const DevConsts *__restrict__ C, int len, SoA y, int incy,
                                                    AoSRef<N> sums) {
    y = rd(y);
    const long long o = group_id<L>();
    if (o >= len) return;
    Grp<N, L> g(C);
    MP<N / L> a, s, r;
    const long long iy = bidx(o, len, incy);
    load_soa(g, a, y, iy);
    load_aos(g, s, sums, o);
    g.add(r, a, s);
    store_soa(g, r, y, iy);
}

// ---------------------------------------------------------------- gemm
// Entry e = (row, col) of the product, column-major: sum_i A[row, i] * B[i, col] in
// order i = 0 .. k-1 (src/blas/gemm.cuh:39-58; v2 src/blas/v2/gemm_v2.cuh:55-75 with
// the transposed index forms). ta/tb are ignored by the v1 layout (SoA).
template <int N, int L, bool FULL, class VM>
__device__ __forceinline__ bool gemm_sum(const Grp<N, L> &g, MP<N / L> &sum, int ta, int tb, int k, const VM &A,
                                         int lda, const VM &B, int ldb, int row, int col) {
    g.zero(sum);
    for (int i = 0; i < k; i++) {
        const long long ia = ta ? (long long)i + (long long)lda * row : (long long)row + (long long)lda * i;
        const long long ib = tb ? (long long)col + (long long)ldb * i : (long long)i + (long long)ldb * col;
        if (!g.template fma_step<FULL>(sum, A, ia, B, ib)) return false;
    }
    return true;
}
// sums of all m*n entries into the AoS workspace. FULL = false, FIX = false: the
// fast kernel (REDO marks an abandoned chain); FULL = true, FIX = true: the fixup
// kernel, recomputing the marked sums only; FULL = true, FIX = false: one full pass.
template <int N, int L, bool FULL, bool FIX, class VM>
__global__ __launch_bounds__(BLK) void k_gemm_sum(const DevConsts *__restrict__ C_, int ta, int tb, int m, int n,
                                                  int k, VM A, int lda, VM B, int ldb, AoSRef<N> sums) {
    if constexpr (std::is_same<VM, SoA>::value) {
        A = rd(A);
        B = rd(B);
    }
    const long long e = group_id<L>();
    if (e >= (long long)m * n) return;
    if (FIX && *sums.sign(e) != REDO) return;
    const int row = (int)(e % m), col = (int)(e / m);
    Grp<N, L> g(C_);
    MP<N / L> sum;
    if (gemm_sum<N, L, FULL>(g, sum, ta, tb, k, A, lda, B, ldb, row, col))
        store_aos(g, sum, sums, e);
    else if (g.lane() == 0)
        *sums.sign(e) = REDO;
}
// Epilogue per entry over the finished sums.
// v1 (src/blas/gemm.cuh:127-166): buffer = sum*alpha (+round), C = C*beta (+round),
//     C = C (+) buffer with the v1 matrix add (+round)
// v2 (src/blas/v2/gemm_v2.cuh:76-78): C = (sum*alpha) (+) (C*beta)
template <int N, int L, bool V2, class VM>
__global__ __launch_bounds__(BLK) void k_gemm_fin(const DevConsts *__restrict__ C_, int m, int n, VM alpha, VM beta,
                                                  VM Cm, int ldc, AoSRef<N> sums) {
    if constexpr (!V2) {
        alpha = rd(alpha);
        beta = rd(beta);
        Cm = rd(Cm);
    }
    const long long e = group_id<L>();
    if (e >= (long long)m * n) return;
    const int row = (int)(e % m), col = (int)(e / m);
    const long long ic = (long long)col * ldc + row;
    Grp<N, L> g(C_);
    MP<N / L> sum, a, b, cv, s2, c2, r;
    load_aos(g, sum, sums, e);
    if constexpr (V2) {
        load_aos(g, a, alpha, 0);
        load_aos(g, b, beta, 0);
        load_aos(g, cv, Cm, ic);
    } else {
        load_soa(g, a, alpha, 0);
        load_soa(g, b, beta, 0);
        load_soa(g, cv, Cm, ic);
    }
    g.mul(s2, sum, a);
    g.mul(c2, cv, b);
    if constexpr (V2) {
        g.add(r, s2, c2);
        store_aos(g, r, Cm, ic);
    } else {
        g.vadd(r, c2, s2);
        store_soa(g, r, Cm, ic);
    }
}

// ---------------------------------------------------------------- gemm, split form
// Over 16 and 32 moduli the sums of gemm are formed in three parts. In a chain
// whose every step is exact (no rounding, no straddling interval, no dropped
// operand) the final residues depend only on the terms, the final exponent E
// and the final sign:
//   digits_r = |sum_i s_i X_i 2^(e_i - E)| mod m_r,
// so the chain runs on the scalar fields alone (k_gemm_chain), and the residues
// come from N independent modular matrix products of pre-scaled operands
//   A'_r[i][row] = s_A x_r 2^(e_A - EA[row]) mod m_r,  B'_r[i][col] likewise,
//   digits_r = 2^(EA[row] + EB[col] - E) * sum_i A'_r[i][row] B'_r[i][col] mod m_r,
// EA / EB being the minimum exponents of each row of op(A) and column of op(B).
// The products run exactly on INT8 MFMA over 7-bit limbs of A' and B'
// (k_gemm_limbs, k_mf_gemm), and k_gemm_digits applies the rescaling. Chains that
// meet a rare case are marked REDO and recomputed step by step by the fixup kernel.


// index of op(A)[row, i] (SIDE 0) or op(B)[i, col] (SIDE 1) in column-major storage
template <int SIDE>
__device__ __forceinline__ long long op_index(int t, int j, int i, int ld) {
    if (SIDE == 0) return t ? (long long)i + (long long)ld * j : (long long)j + (long long)ld * i;
    return t ? (long long)j + (long long)ld * i : (long long)i + (long long)ld * j;
}

// 2^d mod m_r for any integer d (m_r odd)
__device__ __forceinline__ int pow2mod(const DevConsts *C, int d, int r) {
    const double md = C->md[r], mv = C->minv[r];
    const int n = C->n, top = C->log2m;
    if (d >= 0 && d <= top) return C->pow2[(size_t)d * n + r];
    int f = 1;
    if (d >= 0) {
        while (d > top) {
            f = mulmod(f, C->pow2[(size_t)top * n + r], md, mv);
            d -= top;
        }
        return mulmod(f, C->pow2[(size_t)d * n + r], md, mv);
    }
    int D = -d;
    while (D > P2T) {
        f = mulmod(f, C->pow2inv[P2T - 1][r], md, mv);
        D -= P2T;
    }
    return mulmod(f, C->pow2inv[D - 1][r], md, mv);
}

// minimum exponent of each of the cnt rows of op(A) (SIDE 0) or columns of op(B) (SIDE 1),
// over k: workgroup (x, y) covers EMIN_J consecutive j and the k-chunk y of EMIN_K terms,
// combined by atomic minimum into out (filled with INT_MAX by k_fill_int first)
constexpr int EMIN_J = 64, EMIN_K = 64;
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_fill_int(int *p, int cnt, int v) {
    const long long i = (long long)blockIdx.x * BLK + threadIdx.x;
    if (i < cnt) p[i] = v;
}
// SKIPZ: ignore zero elements (interval upper bound 0)
template <int N, int L, class VM, int SIDE, bool SKIPZ = false>
__global__ __launch_bounds__(BLK) void k_gemm_emin(int cnt, int k, int t, VM X, int ld, int *out) {
    if constexpr (std::is_same<VM, SoA>::value) X = rd(X);
    constexpr int S = BLK / EMIN_J;
    const long long j = (long long)blockIdx.x * EMIN_J + threadIdx.x % EMIN_J;
    const int i0 = blockIdx.y * EMIN_K, i1 = (i0 + EMIN_K) < k ? i0 + EMIN_K : k;
    if (j >= cnt) return;
    int e = 0x7fffffff;
    for (int i = i0 + (int)threadIdx.x / EMIN_J; i < i1; i += S) {
        const long long idx = op_index<SIDE>(t, (int)j, i, ld);
        if constexpr (SKIPZ) {
            int s_, e_;
            Er lo_, up_;
            scalars_of<N>(X, idx, s_, e_, lo_, up_);
            if (up_.frac == 0) continue;
        }
        const int v = exp_of<N>(X, idx);
        e = v < e ? v : e;
    }
    if (e != 0x7fffffff) atomicMin(out + j, e);
}

// pre-scaled residue planes P[r][i][j] (j over the cnt rows of op(A) or columns of op(B))
template <int N, int L, class VM, int SIDE>
__global__ __launch_bounds__(BLK) void k_gemm_prescale(const DevConsts *__restrict__ C, int cnt, int k, int t, VM X,
                                                       int ld, const int *__restrict__ emin, int *__restrict__ P) {
    if constexpr (std::is_same<VM, SoA>::value) X = rd(X);
    const long long g = (long long)blockIdx.x * BLK + threadIdx.x;
    if (g >= (long long)cnt * k) return;
    const int j = (int)(g % cnt), i = (int)(g / cnt);
    const long long idx = op_index<SIDE>(t, j, i, ld);
    const int d = exp_of<N>(X, idx) - emin[j];
    const int s = sign_at<N>(X, idx);
    const int *dig = digits_of<N>(X, idx);
    for (int r = 0; r < N; r++) {
        const int m = C->m[r];
        int v = mulmod(dig[r], pow2mod(C, d, r), C->md[r], C->minv[r]);
        v = (s && v) ? m - v : v;
        P[((size_t)r * k + i) * cnt + j] = v;
    }
}

// Residue planes of the split-form gemm for the INT8 MFMA residue products
// (k_mf_gemm): the values of k_gemm_prescale, s x_r 2^(e - emin[j]) mod m_r (signed
// as m_r - v), split into LIMB_BITS-bit limbs, as int8 planes P[r][limb] with j < JP
// (rows of op(A) or columns of op(B), padded with zeros to a multiple of 64) and kp < KP
// (k padded with zeros to a multiple of LIMB_T). A plane is k-step major,
// [kp / LIMB_T][j][kp % LIMB_T], so that one k-step of a panel of rows is one contiguous
// run. k_gemm_limbs: thread g = j * KP + kp reads element (j, kp) (consecutive threads,
// consecutive k) and writes plane position limb_pos(j, kp, JP); k_mf_prescale: thread g
// writes plane position g, element limb_elem(g).
constexpr int LIMB_BITS = 7, LIMB_T = 32;
__device__ __forceinline__ long long limb_pos(int j, int kp, int JP) {
    return ((long long)(kp / LIMB_T) * JP + j) * LIMB_T + kp % LIMB_T;
}
__device__ __forceinline__ void limb_elem(long long g, int JP, int &j, int &kp) {
    const long long rest = g / LIMB_T;
    j = (int)(rest % JP);
    kp = (int)(rest / JP) * LIMB_T + (int)(g % LIMB_T);
}
template <int N, int L, class VM, int SIDE>
__global__ __launch_bounds__(BLK) void k_gemm_limbs(const DevConsts *__restrict__ C, int cnt, int JP, int k, int KP,
                                                    int nl, int t, VM X, int ld, const int *__restrict__ emin,
                                                    signed char *__restrict__ P) {
    if constexpr (std::is_same<VM, SoA>::value) X = rd(X);
    const long long g = (long long)blockIdx.x * BLK + threadIdx.x;
    if (g >= (long long)JP * KP) return;
    const int kp = (int)(g % KP), j = (int)(g / KP);
    const long long pos = limb_pos(j, kp, JP);
    const size_t plane = (size_t)JP * KP;
    if (j >= cnt || kp >= k) {
        for (int r = 0; r < N; r++)
            for (int a = 0; a < nl; a++) P[((size_t)r * nl + a) * plane + pos] = 0;
        return;
    }
    const long long idx = op_index<SIDE>(t, j, kp, ld);
    const int d = exp_of<N>(X, idx) - emin[j];
    const int s = sign_at<N>(X, idx);
    // the element's residues at once (SoA records are 16-byte aligned; AoS 8-byte)
    int dig[N];
    if constexpr (std::is_same<VM, SoA>::value) {
        const int4 *p4 = (const int4 *)digits_of<N>(X, idx);
#pragma unroll
        for (int q = 0; q < N / 4; q++) {
            const int4 v4 = p4[q];
            dig[4 * q] = v4.x;
            dig[4 * q + 1] = v4.y;
            dig[4 * q + 2] = v4.z;
            dig[4 * q + 3] = v4.w;
        }
    } else {
        const int2 *p2 = (const int2 *)digits_of<N>(X, idx);
#pragma unroll
        for (int q = 0; q < N / 2; q++) {
            const int2 v2 = p2[q];
            dig[2 * q] = v2.x;
            dig[2 * q + 1] = v2.y;
        }
    }
    for (int r = 0; r < N; r++) {
        const int m = C->m[r];
        int v = dig[r] ? mulmod(dig[r], pow2mod(C, d, r), C->md[r], C->minv[r]) : 0;
        v = (s && v) ? m - v : v;
        for (int a = 0; a < nl; a++)
            P[((size_t)r * nl + a) * plane + pos] = (signed char)((v >> (LIMB_BITS * a)) & ((1 << LIMB_BITS) - 1));
    }
}
// scalar chain of every entry (one thread per entry): the sum's sign, exponent and
// interval evaluation into the workspace, or REDO
template <int N, int L, class VM>
__global__ __launch_bounds__(BLK) void k_gemm_chain(const DevConsts *__restrict__ C, int ta, int tb, int m, int n,
                                                    int k, VM A, int lda, VM B, int ldb, AoSRef<N> sums) {
    if constexpr (std::is_same<VM, SoA>::value) {
        A = rd(A);
        B = rd(B);
    }
    const long long e = (long long)blockIdx.x * BLK + threadIdx.x;
    if (e >= (long long)m * n) return;
    const int row = (int)(e % m), col = (int)(e / m);
    const ScC K = sc_consts(C);
    Sc acc{0, 0, Er{0.0, 0}, Er{0.0, 0}};
    bool ok = true;
    for (int i = 0; i < k && ok; i++) {
        const long long ia = op_index<0>(ta, row, i, lda), ib = op_index<1>(tb, col, i, ldb);
        int xs, xe, ys, ye;
        Er x0, x1, y0, y1;
        scalars_of<N>(A, ia, xs, xe, x0, x1);
        scalars_of<N>(B, ib, ys, ye, y0, y1);
        ok = sc_fma_step(K, acc, xs, xe, x0, x1, ys, ye, y0, y1);
    }
    if (!ok) {
        *sums.sign(e) = REDO;
        return;
    }
    *sums.sign(e) = acc.sign;
    *sums.exp(e) = acc.exp;
    ErMem *ev = sums.eval(e);
    ev[0] = ErMem{acc.ev0.frac, (long long)acc.ev0.exp};
    ev[1] = ErMem{acc.ev1.frac, (long long)acc.ev1.exp};
}

// k_gemm_chain for the v1 layout and op(A) = A, op(B) = B, with every element offset of A
// and B below 2^28 (so byte offsets of the 16-byte evaluations fit 32 bits): each step
// reads its operands at 32-bit offsets from uniform bases, advanced by lda and 1. UB: m is
// a multiple of 64, so each wave's 64 entries share one column and the column's B fields
// are wave-uniform; they are then read through scalar loads and held in scalar registers.
template <int N, int L, bool UB>
__global__ __launch_bounds__(BLK) void k_gemm_chain_nn(const DevConsts *__restrict__ C, int m, int n, int k, SoA A,
                                                       int lda, SoA B, int ldb, AoSRef<N> sums) {
    A = rd(A);
    B = rd(B);
    const long long e = (long long)blockIdx.x * BLK + threadIdx.x;
    if (e >= (long long)m * n) return;
    const int row = (int)(e % m);
    int col = (int)(e / m);
    if constexpr (UB) col = __builtin_amdgcn_readfirstlane(col);
    const ScC K = sc_consts(C);
    const char *__restrict__ as = (const char *)A.s, *__restrict__ ae = (const char *)A.e;
    const char *__restrict__ a0 = (const char *)A.ev, *__restrict__ a1 = (const char *)(A.ev + A.len);
    const int *__restrict__ bs = B.s;
    const int *__restrict__ be = B.e;
    const ErMem *__restrict__ b0 = B.ev;
    const ErMem *__restrict__ b1 = B.ev + B.len;
    Sc acc{0, 0, Er{0.0, 0}, Er{0.0, 0}};
    bool ok = true;
    const unsigned ib0 = (unsigned)col * (unsigned)ldb;
    // the operands of step i + 1 are loaded before step i runs (the last step reloads its
    // own), so no load latency sits on the chain
    struct Opnd {
        int xs, xe, ys, ye;
        ErMem xl, xu, yl, yu;
    };
    auto load = [&](int i) {
        const unsigned ia = (unsigned)row + (unsigned)i * (unsigned)lda;
        const unsigned o4 = ia << 2, o16 = ia << 4, ib = ib0 + (unsigned)i;
        return Opnd{*(const int *)(as + o4), *(const int *)(ae + o4), bs[ib], be[ib],
                    *(const ErMem *)(a0 + o16), *(const ErMem *)(a1 + o16), b0[ib], b1[ib]};
    };
    Opnd cur = load(0);
    for (int i = 0; i < k && ok; i++) {
        const Opnd nxt = load(i + 1 < k ? i + 1 : i);
        ok = sc_fma_step(K, acc, cur.xs, cur.xe, Er{cur.xl.frac, (int)cur.xl.exp}, Er{cur.xu.frac, (int)cur.xu.exp},
                         cur.ys, cur.ye, Er{cur.yl.frac, (int)cur.yl.exp}, Er{cur.yu.frac, (int)cur.yu.exp});
        cur = nxt;
    }
    if (!ok) {
        *sums.sign(e) = REDO;
        return;
    }
    *sums.sign(e) = acc.sign;
    *sums.exp(e) = acc.exp;
    ErMem *ev = sums.eval(e);
    ev[0] = ErMem{acc.ev0.frac, (long long)acc.ev0.exp};
    ev[1] = ErMem{acc.ev1.frac, (long long)acc.ev1.exp};
}

// exact x mod m for an integer-valued 0 <= x < 2^53
__device__ __forceinline__ double fmod_exact(double x, double md, double mv) {
    const double q = floor(x * mv);
    return mod_fix(fma(-q, md, x), md);
}

// digit r of entry e from the modular dot product S (0 <= S < m)
template <int N>
__device__ __forceinline__ void resid_digit(const DevConsts *C, int r, int S, int ea, int eb, AoSRef<N> sums,
                                            long long e) {
    const int sign = *sums.sign(e);
    if (sign == REDO) return;
    const int t = mulmod(S, pow2mod(C, ea + eb - *sums.exp(e), r), C->md[r], C->minv[r]);
    sums.digits(e)[r] = (sign && t) ? C->m[r] - t : t;
}

// digits of every entry from its residue sums S[r][e] (the INT8 MFMA residue products).
// A workgroup's BLK entries are consecutive records; their digits are formed in
// LDS and written out so that consecutive lanes store consecutive words.
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_gemm_digits(const DevConsts *__restrict__ C, int m, int n,
                                                     const int *__restrict__ S, const int *__restrict__ EA,
                                                     const int *__restrict__ EB, AoSRef<N> sums) {
    __shared__ int sd[BLK * (N + 1)];
    __shared__ int ss[BLK];
    const long long e0 = (long long)blockIdx.x * BLK, e = e0 + threadIdx.x;
    const long long mn = (long long)m * n;
    const int cnt = (int)((mn - e0) < BLK ? (mn - e0) : BLK);
    int sign = REDO;
    if (e < mn) {
        const int row = (int)(e % m), col = (int)(e / m);
        sign = *sums.sign(e);
        if (sign != REDO) {
            const int d = EA[row] + EB[col] - *sums.exp(e);
            for (int r = 0; r < N; r++) {
                const int t = mulmod(S[(size_t)r * mn + e], pow2mod(C, d, r), C->md[r], C->minv[r]);
                sd[threadIdx.x * (N + 1) + r] = (sign && t) ? C->m[r] - t : t;
            }
        }
    }
    ss[threadIdx.x] = sign;
    __syncthreads();
    for (int q = threadIdx.x; q < cnt * N; q += BLK) {
        const int i = q / N, r = q % N;
        if (ss[i] != REDO) sums.digits(e0 + i)[r] = sd[i * (N + 1) + r];
    }
}

// ---------------------------------------------------------------- gemv, split form
// Main pass without transpose over 16 and 32 moduli. As for gemm, an exact chain's
// residues are |sum_j s_j X_j 2^(e_j - E)| mod m, where the sign, the exponent E and
// the interval evaluation come from the interval chain on scalars alone. So the
// partial (o, t) is formed by two kernels: k_gemv_chain runs the chain
// (sc_fma_step) over A's scalar fields and writes sign, exponent and evaluation,
// or REDO; k_gemv_resid then accumulates the residues against the fixed exponent
// E_ref = EA[o] + EB (EA[o]: minimum exponent of row o of A, EB: minimum exponent
// of bx) and rescales once by 2^(E_ref - E). bx is pre-scaled
// (bxp[r][j] = s x_r 2^(e_j - EB) mod m_r); per step the workgroup builds in LDS
// the 64 multiples bxp[r][j] 2^d (d < 64), so a term costs one product mod m per
// residue; d >= 64 takes the slow path. Chains marked REDO are left to the fixup
// kernel.
constexpr int GV_D = 64;
// scalar chain of the partial (o, t), o fastest (one thread each, so that a
// wavefront reads a run of one column of A's scalar fields per step)
// NARROW: every element offset of A is below 2^28, so the 16-byte evaluations are read
// at 32-bit byte offsets from uniform bases. UB: m is a multiple of 64, so each wave's
// 64 chains share t and read the same bx entry; it then comes through scalar loads.
// The operands of step j + b3 are loaded before step j runs (the last step reloads its
// own), so the chain does not wait on memory.
template <int N, int L, bool NARROW, bool UB>
__global__ __launch_bounds__(BLK) void k_gemv_chain(const DevConsts *__restrict__ C, int m, int n, SoA A, int lda,
                                                    AoSRef<N> bx, int b3, AoSRef<N> part) {
    A = rd(A);
    const long long g = (long long)blockIdx.x * BLK + threadIdx.x;
    if (g >= (long long)m * b3) return;
    const int o = (int)(g % m);
    int t = (int)(g / m);
    if constexpr (UB) t = __builtin_amdgcn_readfirstlane(t);
    const ScC K = sc_consts(C);
    const char *__restrict__ as = (const char *)A.s, *__restrict__ ae = (const char *)A.e;
    const char *__restrict__ a0 = (const char *)A.ev, *__restrict__ a1 = (const char *)(A.ev + A.len);
    const char *__restrict__ xb = bx.base;
    struct Opnd {
        int xs, xe, ys, ye;
        ErMem xl, xu, yl, yu;
    };
    auto load = [&](int j) {
        using Off = typename std::conditional<NARROW, unsigned, unsigned long long>::type;
        const Off ia = (Off)o + (Off)j * (Off)lda, o4 = ia << 2, o16 = ia << 4;
        const char *y = xb + (size_t)j * AoSRef<N>::STRIDE;
        return Opnd{*(const int *)(as + o4), *(const int *)(ae + o4), *(const int *)(y + 4 * N),
                    *(const int *)(y + 4 * N + 4), *(const ErMem *)(a0 + o16), *(const ErMem *)(a1 + o16),
                    *(const ErMem *)(y + 4 * N + 8), *(const ErMem *)(y + 4 * N + 24)};
    };
    Sc sc{0, 0, Er{0.0, 0}, Er{0.0, 0}};
    bool ok = true;
    if (t < n) {
        Opnd cur = load(t);
        for (int j = t; j < n && ok; j += b3) {
            const Opnd nxt = load(j + b3 < n ? j + b3 : j);
            ok = sc_fma_step(K, sc, cur.xs, cur.xe, Er{cur.xl.frac, (int)cur.xl.exp}, Er{cur.xu.frac, (int)cur.xu.exp},
                             cur.ys, cur.ye, Er{cur.yl.frac, (int)cur.yl.exp}, Er{cur.yu.frac, (int)cur.yu.exp});
            cur = nxt;
        }
    }
    const long long v = (long long)o * b3 + t;
    if (!ok) {
        *part.sign(v) = REDO;
        return;
    }
    *part.sign(v) = sc.sign;
    *part.exp(v) = sc.exp;
    ErMem *ev = part.eval(v);
    ev[0] = ErMem{sc.ev0.frac, (long long)sc.ev0.exp};
    ev[1] = ErMem{sc.ev1.frac, (long long)sc.ev1.exp};
}

constexpr int GV_REDUCE = 1 << 20;  // terms between reductions of the double accumulators
// The residues of the partials, aligned at the row's minimum exponent EA[o] + EB[0]: the
// residues of the exact signed sum 2^-(EA[o] + EB[0]) sum_j A[o, j] bx[j]. They do not
// depend on the scalar chains (k_gemv_chain), so the two kernels may run at the same
// time; k_gemv_scale then moves them to the chain's exponent and sign. The N residues of
// a row are spread over N / GR threads (GR residues each) so that m * b3 * N / GR
// threads hide the memory latency that one thread per row cannot at these grid sizes. Thread q of a workgroup holds
// residues (q % TPR) * GR .. + GR - 1 of row q / TPR: a wave's two 16-byte loads per
// step cover a contiguous run of whole rows of column j, so A's residues need no LDS
// staging. The multiples table (bxp 2^d, d < 64) is double-buffered in LDS: one
// barrier per step. Rare terms (d >= 64) first fold 2^(d - d mod 64) into the
// thread's own residues.
constexpr int GR = 8, GV_RT = 512;  // residues per thread, threads per workgroup of k_gemv_resid
template <int N, int L>
__global__ __launch_bounds__(GV_RT) void k_gemv_resid(const DevConsts *__restrict__ C, int m, int n, SoA A, int lda,
                                                      const int *__restrict__ bxp, const int *__restrict__ EA,
                                                      const int *__restrict__ EB, int b3, AoSRef<N> part) {
    static_assert(N % GR == 0 && (GV_D * N) % GV_RT == 0, "residue groups");
    constexpr int TPR = N / GR, RW = GV_RT / TPR, SP = N + 1, DB = GV_D * N / GV_RT;
    __shared__ int tp[2][GV_D * SP];
    A = rd(A);
    const int t = blockIdx.y;
    const int g0 = (threadIdx.x % TPR) * GR;
    const long long o = (long long)blockIdx.x * RW + threadIdx.x / TPR;
    const bool ok = o < m;
    const long long v = o * b3 + t;
    const int ea = ok ? EA[o] : 0;
    double md[GR], mv[GR], acc[GR];
#pragma unroll
    for (int r = 0; r < GR; r++) {
        md[r] = C->md[g0 + r];
        mv[r] = C->minv[g0 + r];
        acc[r] = 0.0;
    }
    int since = 0;
    // table build: thread q -> residue q % N, multiples d0 .. d0 + DB - 1
    const int tr = threadIdx.x % N, td0 = (threadIdx.x / N) * DB;
    const double tmd = C->md[tr], tmv = C->minv[tr];
    const int tm = C->m[tr];
    const int t2 = C->pow2[(size_t)td0 * N + tr];
    // step j's table seed, sign, exponent and residues are loaded during step j - b3
    int nb = 0, nxs = 0, nxe = 0;
    int4 nd0 = make_int4(0, 0, 0, 0), nd1 = nd0;
    auto fetch = [&](int j) {
        nb = bxp[(size_t)tr * n + j];
        if (ok) {
            const long long ia = o + (long long)j * lda;
            nxs = A.s[ia];
            nxe = A.e[ia];
            const int4 *p = (const int4 *)(A.d + ia * N + g0);
            nd0 = p[0];
            nd1 = p[1];
        }
    };
    if (t < n) fetch(t);
    int buf = 0;
    for (int j = t; j < n; j += b3, buf ^= 1) {
        int *tb = tp[buf];
        int tv = mulmod(nb, t2, tmd, tmv);
#pragma unroll
        for (int d = 0; d < DB; d++) {
            tb[(td0 + d) * SP + tr] = tv;
            tv += tv;
            tv = tv >= tm ? tv - tm : tv;
        }
        const int xs = nxs, xe = nxe;
        int xd[GR] = {nd0.x, nd0.y, nd0.z, nd0.w, nd1.x, nd1.y, nd1.z, nd1.w};
        if (j + b3 < n) fetch(j + b3);
        __syncthreads();
        if (ok) {
            int d = xe - ea;
            if (d >= GV_D) {
                const int dh = d - (d & (GV_D - 1));
#pragma unroll
                for (int r = 0; r < GR; r++) xd[r] = mulmod(xd[r], pow2mod(C, dh, g0 + r), md[r], mv[r]);
                d &= GV_D - 1;
            }
            const double sg = xs ? -1.0 : 1.0;
            const int *f = tb + d * SP + g0;
#pragma unroll
            for (int r = 0; r < GR; r++) {
                const double x = (double)xd[r], fr = (double)f[r];
                const double p = x * fr, pl = fma(x, fr, -p);
                const double q = floor(p * mv[r]);
                acc[r] = fma(sg, fma(-q, md[r], p) + pl, acc[r]);
            }
            if (++since == GV_REDUCE) {
                since = 0;
#pragma unroll
                for (int r = 0; r < GR; r++) acc[r] = fmod_exact(acc[r], md[r], mv[r]);
            }
        }
    }
    if (!ok) return;
    int *dig = part.digits(v) + g0;
#pragma unroll
    for (int r = 0; r < GR; r++) dig[r] = (int)fmod_exact(acc[r], md[r], mv[r]);
}
// The residues of k_gemv_resid (the signed sum aligned at EA[o] + EB[0]) moved to the
// chain's result: times 2^(EA[o] + EB[0] - exp), and m - w for a negative sum (digits
// hold the magnitude). Partials whose chains were abandoned (REDO) are left to the fixup,
// which runs after this kernel. GR residues per thread; nvt = m * b3 partials.
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_gemv_scale(const DevConsts *__restrict__ C, long long nvt, int b3,
                                                    const int *__restrict__ EA, const int *__restrict__ EB,
                                                    AoSRef<N> part) {
    constexpr int TPR = N / GR;
    const long long q = (long long)blockIdx.x * BLK + threadIdx.x;
    const long long v = q / TPR;
    if (v >= nvt) return;
    const int g0 = (int)(q % TPR) * GR;
    const int csign = *part.sign(v);
    if (csign == REDO) return;
    const int dz = EA[v / b3] + EB[0] - *part.exp(v);
    int2 *p = (int2 *)(part.digits(v) + g0);  // records are 8-byte aligned
#pragma unroll
    for (int h = 0; h < GR / 2; h++) {
        const int2 x = p[h];
        int w[2] = {x.x, x.y};
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int r = g0 + 2 * h + u;
            w[u] = mulmod(w[u], pow2mod(C, dz, r), C->md[r], C->minv[r]);
            w[u] = (csign && w[u]) ? C->m[r] - w[u] : w[u];
        }
        p[h] = make_int2(w[0], w[1]);
    }
}

// ---------------------------------------------------------------- SpMV
// Row partial sums of tpr strided lanes (src/sparse/spmv/spmv_csrv.cuh:61-65; tpr = 1
// is the scalar kernel spmv_csr.cuh:53-62), combined by the 16/8/4/2/1 tree
// (:66-80) through LDS. One workgroup holds BLK/L groups = (BLK/L)/tpr rows.
// The fast kernel marks a row whose chain it abandoned with y[row].sign = REDO;
// the fixup kernel recomputes those rows only.
//
// The storage formats differ only in which terms a lane visits, in which order, and
// where the row's result goes (FMT):
//   SP_CSR  entries irp[row] + lane, + tpr, ... (spmv_csr.cuh:53-62, spmv_csrv.cuh:61-65)
//   SP_ELL  slots c = lane, + tpr, ... of the column-major m x width arrays, skipping
//           zero values (spmv_ell.cuh:53-66)
//   SP_JAD  jagged diagonals j = lane, + tpr, ... while row + jcp[j] < jcp[j + 1];
//           the result goes to y[perm[row]] (spmv_jad.cuh:53-66, spmv_jadv.cuh:53-62)
//   SP_DIA  diagonals i = lane, + tpr, ... with 0 <= row + offset[i] < ncols and a
//           nonzero value (spmv_dia.cuh:53-66)
template <int FMT>
__device__ __forceinline__ long long sp_out_row(const SpMat &A, long long row) {
    if constexpr (FMT == SP_JAD) return A.perm[row];
    return row;
}
// the row's terms of one lane, in the reference's order: f(column, value) until it returns false
template <int FMT, class F>
__device__ __forceinline__ void sp_terms(const SpMat &A, int m, long long row, int lane, int tpr, F &&f) {
    if constexpr (FMT == SP_CSR) {
        const int re = A.ptr[row + 1];
        for (int i = A.ptr[row] + lane; i < re; i += tpr)
            if (!f(A.ja[i], A.as[i])) return;
    } else if constexpr (FMT == SP_ELL) {
        for (int c = lane; c < A.width; c += tpr) {
            const long long q = (long long)c * m + row;
            const double v = A.as[q];
            if (v != 0 && !f(A.ja[q], v)) return;
        }
    } else if constexpr (FMT == SP_JAD) {
        for (int j = lane; j < A.width; j += tpr) {
            const long long q = row + A.ptr[j];
            if (q < A.ptr[j + 1] && !f(A.ja[q], A.as[q])) return;
        }
    } else {
        for (int i = lane; i < A.width; i += tpr) {
            const long long col = row + A.offset[i];
            const double v = A.as[(long long)m * i + row];
            if (col >= 0 && col < A.ncols && v != 0 && !f((int)col, v)) return;
        }
    }
}
// one row of k_spmv (its tree synchronises the workgroup: every thread calls it)
template <int N, int L, bool FULL, bool FIX, int FMT, int W>
__device__ __forceinline__ void spmv_row(const DevConsts *__restrict__ C, int m, const SpMat &A, AoSRef<N> x,
                                         AoSRef<N> y, int tpr, bool live, long long row, int lane, int gl,
                                         int (&sd)[W][N], int (&ss)[W], int (&se)[W], double (&sf0)[W],
                                         double (&sf1)[W], int (&sx0)[W], int (&sx1)[W]) {
    constexpr int R = N / L;
    const long long orow = live ? sp_out_row<FMT>(A, row) : 0;
    if (FIX && live) live = *y.sign(orow) == REDO;
    Grp<N, L> g(C);
    MP<R> acc, p;
    g.zero(acc);
    bool ok = true;
    if (live)
        sp_terms<FMT>(A, m, row, lane, tpr, [&](int col, double v) {
            ok = g.template fma_d_step<FULL>(acc, x, col, v);
            return ok;
        });
    if (tpr > 1) {
        for (int h = tpr >> 1; h >= 1; h >>= 1) {
            // publish partials of lanes h..2h-1, then lanes < h add them
            if (live && lane >= h && lane < 2 * h) {
#pragma unroll
                for (int r = 0; r < R; r++) sd[gl][g.ix(r)] = acc.d[r];
                if (g.lane() == 0) {
                    ss[gl] = ok ? acc.sign : REDO;
                    se[gl] = acc.exp;
                    sf0[gl] = acc.ev0.frac;
                    sx0[gl] = acc.ev0.exp;
                    sf1[gl] = acc.ev1.frac;
                    sx1[gl] = acc.ev1.exp;
                }
            }
            __syncthreads();
            if (live && lane < h) {
                const int o = gl + h;
#pragma unroll
                for (int r = 0; r < R; r++) p.d[r] = sd[o][g.ix(r)];
                p.sign = ss[o];
                p.exp = se[o];
                p.ev0 = Er{sf0[o], sx0[o]};
                p.ev1 = Er{sf1[o], sx1[o]};
                if (ok && p.sign == REDO) ok = false;
                if (ok) ok = g.template add_try<FULL>(acc, acc, p);
            }
            __syncthreads();
        }
    }
    if (live && lane == 0) {
        if (ok)
            store_aos(g, acc, y, orow);
        else if (g.lane() == 0)
            *y.sign(orow) = REDO;
    }
}

// k_spmv over rows 0 .. m-1, or over a row list (the fixup only: *rcnt rows in rlist, as
// marked by the split form's chain kernel): the workgroups then take list entries blockIdx.x * RPB, + gridDim.x * RPB, ... (RPB
// rows per workgroup), so any grid covers the list.
template <int N, int L, bool FULL, bool FIX, int FMT = SP_CSR>
__global__ __launch_bounds__(BLK) void k_spmv(const DevConsts *__restrict__ C, int m, SpMat A, AoSRef<N> x,
                                              AoSRef<N> y, int tpr, const int *rcnt = nullptr,
                                              const int *rlist = nullptr) {
    constexpr int W = BLK / L;
    __shared__ int sd[W][N];
    __shared__ int ss[W], se[W];
    __shared__ double sf0[W], sf1[W];
    __shared__ int sx0[W], sx1[W];
    const int gl = threadIdx.x / L;
    const int RPB = W / tpr;
    const int lane = gl % tpr;
    const long long count = rlist ? (long long)*rcnt : (long long)m;
    const long long stride = rlist ? (long long)gridDim.x * RPB : count;
    for (long long base = (long long)blockIdx.x * RPB; base < count; base += stride) {
        const long long idx = base + gl / tpr;
        spmv_row<N, L, FULL, FIX, FMT>(C, m, A, x, y, tpr, idx < count, rlist ? (idx < count ? rlist[idx] : 0) : idx,
                                       lane, gl, sd, ss, se, sf0, sf1, sx0, sx1);
    }
}

// ---------------------------------------------------------------- SpMV, split form
// At 16 and 32 moduli a row's exact chains (every product and sum exact, no sign
// ambiguity, no nonzero operand dropped) end in residues |sum_i s_i X_i sig_i 2^(e_i - E)|
// mod m, where term i is x[col] times the double value (significand sig_i, exponent
// e_i of the product), s_i the product's sign and E the row result's exponent: the
// chains and the tree only align and add, and alignment multiplies by powers of two.
// So the row is formed by separate kernels: k_spmv_chain runs the reference's chains
// and tree on the scalar fields alone (one thread per chain, where the fused kernel
// repeats the scalar work on every lane of a number) and writes y's sign, exponent and
// evaluation, or REDO; k_spmv_resid sums the residues of the row's terms against E, in
// any order (the sum is exact), CSR rows of at least 2^SP_LONG_BITS terms by
// k_spmv_resid_wave with SPW_T threads per row (the two residue kernels take disjoint rows
// and may run at once). REDO rows go to the fused kernel's fixup pass.
constexpr int SGR = 4;           // residues per thread in the residue kernels
constexpr int SP_LONG_BITS = 8;  // CSR rows of 2^8 = 128 terms or more take the long-row residue kernel
// the long rows' length classes lead the row order (k_sp_bin_*): off[SP_LONG_CLASS] counts them
constexpr int SP_LONG_CLASS = SP_NB - SP_LONG_BITS;
constexpr int SP_PF = 4;      // terms of a chain whose x fields are in flight ahead of the step
constexpr int SP_PF1 = 8;     // the same for one chain per row (CSR scalar kernel: long serial chains)
#ifndef MPH_SP_TS
#define MPH_SP_TS 2
#endif
constexpr int SP_TS = MPH_SP_TS;  // term slots per row in k_spmv_resid (1 in the one-lane CPU emulation)

// sign, exponent and evaluation of x[col] (40 contiguous bytes of the record)
struct XScal {
    int2 se;
    ErMem l, u;
};
template <int N>
__device__ __forceinline__ XScal load_xscal(const AoSRef<N> &x, int col) {
    const char *p = x.base + (size_t)col * AoSRef<N>::STRIDE + 4 * N;
    return XScal{*(const int2 *)p, *(const ErMem *)(p + 8), *(const ErMem *)(p + 24)};
}

// perm (CSR): rows by decreasing length class (k_sp_bin_*), so that the chains of one
// wavefront have similar lengths; rows marked REDO are appended to rlist (count *rcnt)
template <int N, int FMT, int PF = SP_PF>
__global__ __launch_bounds__(BLK) void k_spmv_chain(const DevConsts *__restrict__ C, int m, SpMat A, AoSRef<N> x,
                                                    AoSRef<N> y, int tpr, const int *__restrict__ perm, int *rcnt,
                                                    int *rlist) {
    const long long g = (long long)blockIdx.x * BLK + threadIdx.x;
    const long long idx = g / tpr;
    const int lane = (int)(g % tpr);
    const bool live = idx < m;
    const long long row = live ? (perm ? perm[idx] : idx) : 0;
    const ScC K = sc_consts(C);
    Sc acc{0, 0, Er{0.0, 0}, Er{0.0, 0}};
    bool ok = true;
    auto step = [&](const XScal &o, double v) {
        return sc_muld_step(K, acc, o.se.x, o.se.y, Er{o.l.frac, (int)o.l.exp}, Er{o.u.frac, (int)o.u.exp}, v);
    };
    if constexpr (FMT == SP_CSR) {
        // terms i, i + tpr, ... in blocks of PF: x's fields of a block are gathered one
        // block ahead, its columns and values two blocks ahead, into fixed registers (a
        // ring that is never shifted, so a step waits only for its own operands)
        if (live) {
            const int re = A.ptr[row + 1];
            int i = A.ptr[row] + lane;
            auto col = [&](int j) { return j < re ? A.ja[j] : 0; };
            auto val = [&](int j) { return j < re ? A.as[j] : 0.0; };
            XScal b[PF];
            double vb[PF], vn[PF];
            int cn[PF];
#pragma unroll
            for (int k = 0; k < PF; k++) {
                b[k] = load_xscal(x, col(i + k * tpr));
                vb[k] = val(i + k * tpr);
                cn[k] = col(i + (PF + k) * tpr);
                vn[k] = val(i + (PF + k) * tpr);
            }
            for (; i < re && ok; i += PF * tpr) {
#pragma unroll
                for (int k = 0; k < PF; k++) {
                    if (ok && i + k * tpr < re) ok = step(b[k], vb[k]);
                    b[k] = load_xscal(x, cn[k]);
                    vb[k] = vn[k];
                    cn[k] = col(i + (2 * PF + k) * tpr);
                    vn[k] = val(i + (2 * PF + k) * tpr);
                }
            }
        }
    } else if (live) {
        sp_terms<FMT>(A, m, row, lane, tpr, [&](int col, double v) {
            ok = step(load_xscal(x, col), v);
            return ok;
        });
    }
    // the 16/8/4/2/1 tree (spmv_csrv.cuh:66-80): lane l < h adds the partial of lane l + h;
    // a row's tpr lanes are adjacent in one wavefront, and every lane takes part in the shuffles
    for (int h = tpr >> 1; h >= 1; h >>= 1) {
        Sc p;
        p.sign = __shfl_down(acc.sign, h, tpr);
        p.exp = __shfl_down(acc.exp, h, tpr);
        p.ev0 = Er{__shfl_down(acc.ev0.frac, h, tpr), __shfl_down(acc.ev0.exp, h, tpr)};
        p.ev1 = Er{__shfl_down(acc.ev1.frac, h, tpr), __shfl_down(acc.ev1.exp, h, tpr)};
        const int pok = __shfl_down((int)ok, h, tpr);
        if (lane < h) {
            ok = ok && pok;
            if (ok) ok = sc_add_step(K, acc, p);
        }
    }
    if (!live || lane != 0) return;
    const long long orow = sp_out_row<FMT>(A, row);
    if (!ok) {
        *y.sign(orow) = REDO;
        rlist[atomicAdd(rcnt, 1)] = (int)row;
        return;
    }
    *y.sign(orow) = acc.sign;
    *y.exp(orow) = acc.exp;
    ErMem *ev = y.eval(orow);
    ev[0] = ErMem{acc.ev0.frac, (long long)acc.ev0.exp};
    ev[1] = ErMem{acc.ev1.frac, (long long)acc.ev1.exp};
}

// the operands a residue thread reads of x[col]: residues g0 .. g0 + SGR - 1, sign, exponent
struct XRes {
    int2 se, a, b;  // sign and exponent of x; residues g0 .. g0 + 3
};
template <int N>
__device__ __forceinline__ XRes load_xres(const AoSRef<N> &x, int col, int g0) {
    const int *d = x.digits(col) + g0;
    return XRes{*(const int2 *)x.sign(col), *(const int2 *)d, *(const int2 *)(d + 2)};
}
// one term's residues from its operands and the power 2^d mod m_r of its alignment: acc_r +=
// s (x_r f_r mod m_r) (sig mod m_r) as an exact remainder
__device__ __forceinline__ void sp_resid_add(const double *md, const double *mv, double *acc, const XRes &o, double v,
                                             const int *f) {
    if (v == 0) return;  // a zero product adds nothing
    const uint64_t u = d2u(v);
    const double sig = (double)((u & ((((uint64_t)1) << 52) - 1)) | (((uint64_t)1) << 52));
    const double sg = (o.se.x ^ (int)(u >> 63)) ? -1.0 : 1.0;
    const int xd[SGR] = {o.a.x, o.a.y, o.b.x, o.b.y};
#pragma unroll
    for (int r = 0; r < SGR; r++) {
        const double a = (double)mulmod(xd[r], f[r], md[r], mv[r]), b = (double)modd(sig, md[r], mv[r]);
        const double pr = a * b, pl = fma(a, b, -pr);
        const double q = floor(pr * mv[r]);
        acc[r] = fma(sg, fma(-q, md[r], pr) + pl, acc[r]);
    }
}
// d = e - E of a term: 0 <= d <= log2 M for every nonzero term of an exact chain (its aligned
// value is below M); a zero x may fall outside, adds nothing, and is clamped into the table
__device__ __forceinline__ int sp_resid_shift(const XRes &o, double v, int E, int top) {
    const int d = o.se.y + ((int)((d2u(v) >> 52) & 2047) - 1075) - E;
    return d < 0 ? 0 : (d > top ? top : d);
}
// terms i0, i0 + ts, ... below re of a CSR row (rs: a valid index of the row) added to acc,
// in blocks of PF: a block's table entries are requested before the next block's operands,
// which are gathered while the block is added
template <int N, int PF>
__device__ __forceinline__ void sp_resid_slot(const DevConsts *__restrict__ C, const SpMat &A, const AoSRef<N> &x,
                                              int g0, const double *md, const double *mv, double *acc, int i0, int re,
                                              int rs, int ts, int E) {
    const int top = C->log2m;
    const int *__restrict__ p2 = C->pow2 + g0;
    auto ld = [&](int j) { return load_xres(x, A.ja[j < re ? j : rs], g0); };
    auto lv = [&](int j) { return j < re ? A.as[j] : 0.0; };
    XRes b[PF];
    double vb[PF];
#pragma unroll
    for (int k = 0; k < PF; k++) {
        b[k] = ld(i0 + k * ts);
        vb[k] = lv(i0 + k * ts);
    }
    for (int i = i0; i < re; i += PF * ts) {
        int f[PF][SGR];
#pragma unroll
        for (int k = 0; k < PF; k++) {
            const int2 *t = (const int2 *)(p2 + (size_t)sp_resid_shift(b[k], vb[k], E, top) * N);
            const int2 t0 = t[0], t1 = t[1];
            f[k][0] = t0.x;
            f[k][1] = t0.y;
            f[k][2] = t1.x;
            f[k][3] = t1.y;
        }
        XRes nb[PF];
        double nv[PF];
#pragma unroll
        for (int k = 0; k < PF; k++) {
            nb[k] = ld(i + (PF + k) * ts);
            nv[k] = lv(i + (PF + k) * ts);
        }
#pragma unroll
        for (int k = 0; k < PF; k++)
            if (i + k * ts < re) sp_resid_add(md, mv, acc, b[k], vb[k], f[k]);
#pragma unroll
        for (int k = 0; k < PF; k++) {
            b[k] = nb[k];
            vb[k] = nv[k];
        }
    }
}

// Residues of the rows k_spmv_chain finished: SGR residues per thread, N / SGR residue
// groups by SP_TS term slots per row (slot t takes terms t, t + SP_TS, ...; the slots'
// sums meet through shuffles), rows in perm's order (CSR). With skip (CSR), the rows from
// perm[*skip] on: the first *skip, those of 2^SP_LONG_BITS terms or more, are
// k_spmv_resid_wave's.
template <int N, int FMT>
__global__ __launch_bounds__(BLK) void k_spmv_resid(const DevConsts *__restrict__ C, int m, SpMat A, AoSRef<N> x,
                                                    AoSRef<N> y, const int *__restrict__ perm,
                                                    const int *__restrict__ skip) {
    static_assert(N % SGR == 0 && SGR == 4, "residue groups");
    constexpr int G = N / SGR, TPR = G * SP_TS;
    static_assert(64 % TPR == 0, "a row's threads in one wavefront");
    const long long g = (long long)blockIdx.x * BLK + threadIdx.x;
    const long long idx = g / TPR + (skip ? *skip : 0);
    const int q = (int)(g % G), sl = (int)((g / G) % SP_TS), g0 = q * SGR;
    bool live = idx < m;
    const long long row = live ? (perm ? perm[idx] : idx) : 0;
    const long long orow = live ? sp_out_row<FMT>(A, row) : 0;
    const int ysign = live ? *y.sign(orow) : REDO;
    live = live && ysign != REDO;
    double md[SGR], mv[SGR], acc[SGR];
#pragma unroll
    for (int r = 0; r < SGR; r++) {
        md[r] = C->md[g0 + r];
        mv[r] = C->minv[g0 + r];
        acc[r] = 0.0;
    }
    if constexpr (FMT == SP_CSR) {
        const int rs = live ? A.ptr[row] : 0, re = live ? A.ptr[row + 1] : 0;
        // fewer than 2^SP_LONG_BITS terms with skip (fewer than 2^20 in any case: no
        // intermediate reduction); the next two terms' operands are in flight while one is added
        if (live && rs + sl < re) {  // a slot with terms: rs is then a valid index
            const int E = *y.exp(orow);
            sp_resid_slot<N, 2>(C, A, x, g0, md, mv, acc, rs + sl, re, rs, SP_TS, E);
        }
    } else if (live) {
        const int E = *y.exp(orow);
        int since = 0, t = 0;
        sp_terms<FMT>(A, m, row, 0, 1, [&](int col, double v) {
            if (t++ % SP_TS != sl) return true;
            const XRes o = load_xres(x, col, g0);
            const int2 *t = (const int2 *)(C->pow2 + (size_t)sp_resid_shift(o, v, E, C->log2m) * N + g0);
            const int f[SGR] = {t[0].x, t[0].y, t[1].x, t[1].y};
            sp_resid_add(md, mv, acc, o, v, f);
            if (++since == GV_REDUCE) {
                since = 0;
#pragma unroll
                for (int r = 0; r < SGR; r++) acc[r] = fmod_exact(acc[r], md[r], mv[r]);
            }
            return true;
        });
    }
    // the slots' sums (reduced below m, so SP_TS of them add exactly) through shuffles
    double w[SGR];
#pragma unroll
    for (int r = 0; r < SGR; r++) w[r] = fmod_exact(acc[r], md[r], mv[r]);
#pragma unroll
    for (int h = G; h < TPR; h <<= 1)
#pragma unroll
        for (int r = 0; r < SGR; r++) w[r] += __shfl_xor(w[r], h, 64);
    if (!live || sl != 0) return;
    int *dig = y.digits(orow) + g0;
#pragma unroll
    for (int r = 0; r < SGR; r++) {
        const int v = (int)fmod_exact(w[r], md[r], mv[r]);
        dig[r] = (ysign && v) ? C->m[g0 + r] - v : v;
    }
}

// Residues of the long CSR rows (2^SP_LONG_BITS terms or more: the first *nlong rows of the
// row order perm), a group of SPW_T threads per row: SPW_T / (N / SGR) term slots of N / SGR
// threads, whose reduced sums meet through shuffles (the group is one wavefront, so no
// barrier). Groups take rows g, g + (groups of the grid), ... of the list, so any grid covers
// it; rows the chains marked REDO are left to the fixup pass. (A workgroup per row, 64 slots
// joined in LDS, left most slots idle on rows of a few hundred terms.)
#ifndef MPH_SPW_T
#define MPH_SPW_T 32
#endif
constexpr int SPW_T = MPH_SPW_T;  // threads per long row (half a wavefront; 1 slot per group on the host)
template <int N>
__global__ __launch_bounds__(BLK) void k_spmv_resid_wave(const DevConsts *__restrict__ C, SpMat A, AoSRef<N> x,
                                                         AoSRef<N> y, const int *__restrict__ perm,
                                                         const int *__restrict__ nlong) {
    constexpr int G = N / SGR, TPR = SPW_T < G ? G : SPW_T, TS = TPR / G;
    static_assert(64 % TPR == 0 && BLK % TPR == 0, "a row's threads in one wavefront");
    const int t = threadIdx.x % TPR, q = t % G, sl = t / G, g0 = q * SGR;
    const int grp = blockIdx.x * (BLK / TPR) + threadIdx.x / TPR, ngrp = gridDim.x * (BLK / TPR);
    double md[SGR], mv[SGR];
#pragma unroll
    for (int r = 0; r < SGR; r++) {
        md[r] = C->md[g0 + r];
        mv[r] = C->minv[g0 + r];
    }
    const int count = *nlong;
    for (int k = grp; k < count; k += ngrp) {  // uniform over the row's group
        const int row = perm[k];
        if (*y.sign(row) == REDO) continue;
        const int E = *y.exp(row);
        const int rs = A.ptr[row], re = A.ptr[row + 1];
        double acc[SGR] = {0.0, 0.0, 0.0, 0.0};
        // a slot's terms in blocks of at most GV_REDUCE, reduced between blocks
        for (int i = rs + sl; i < re; i += GV_REDUCE * TS) {
            const long long e = (long long)i + (long long)GV_REDUCE * TS;
            sp_resid_slot<N, 4>(C, A, x, g0, md, mv, acc, i, e < re ? (int)e : re, rs, TS, E);
#pragma unroll
            for (int r = 0; r < SGR; r++) acc[r] = fmod_exact(acc[r], md[r], mv[r]);
        }
        // the slots' sums (each below m, TS of them add exactly) through shuffles
        double w[SGR];
#pragma unroll
        for (int r = 0; r < SGR; r++) w[r] = fmod_exact(acc[r], md[r], mv[r]);
#pragma unroll
        for (int h = G; h < TPR; h <<= 1)
#pragma unroll
            for (int r = 0; r < SGR; r++) w[r] += __shfl_xor(w[r], h, 64);
        if (sl == 0) {
            const int sg = *y.sign(row);
            int *dig = y.digits(row) + g0;
#pragma unroll
            for (int r = 0; r < SGR; r++) {
                const int v = (int)fmod_exact(w[r], md[r], mv[r]);
                dig[r] = (sg && v) ? C->m[g0 + r] - v : v;
            }
        }
    }
}

// ---- rows of a CSR matrix grouped by length class, longest first: class c holds rows of
// 2^(SP_NB - 2 - c) .. 2^(SP_NB - 1 - c) - 1 terms (class 0: all longer, class SP_NB - 1:
// empty rows). A workgroup counts its rows per class in LDS and reserves its ranges
// with one atomic per class; order within a class is arbitrary (each row's own terms
// keep the reference's order).
__device__ __forceinline__ int sp_class(int len) {
    const int b = len > 0 ? 32 - __clz(len) : 0;  // bit length
    const int c = SP_NB - 1 - b;
    return c < 0 ? 0 : c;
}
// hist[SP_NB]: rows per class. A workgroup covers SP_RPT * BLK rows (thread t: rows
// t, t + BLK, ...), so that few workgroups contend for the global counters.
constexpr int SP_RPT = 16;
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_sp_bin_count(int m, const int *__restrict__ irp, int *hist) {
    __shared__ int h[SP_NB];
    if (threadIdx.x < SP_NB) h[threadIdx.x] = 0;
    __syncthreads();
    const long long r0 = (long long)blockIdx.x * BLK * SP_RPT + threadIdx.x;
    for (int k = 0; k < SP_RPT; k++) {
        const long long row = r0 + (long long)k * BLK;
        if (row < m) atomicAdd(&h[sp_class(irp[row + 1] - irp[row])], 1);
    }
    __syncthreads();
    if (threadIdx.x < SP_NB && h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}
// off[c] = rows of the classes before c (one thread)
template <int N, int L>
__global__ void k_sp_bin_scan(const int *hist, int *off) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int a = 0;
    for (int c = 0; c < SP_NB; c++) {
        off[c] = a;
        a += hist[c];
    }
}
// perm[off[c] + cursor] = row; cur[SP_NB] starts at zero
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_sp_bin_scatter(int m, const int *__restrict__ irp, const int *off, int *cur,
                                                        int *perm) {
    __shared__ int h[SP_NB], base[SP_NB];
    if (threadIdx.x < SP_NB) h[threadIdx.x] = 0;
    __syncthreads();
    const long long r0 = (long long)blockIdx.x * BLK * SP_RPT + threadIdx.x;
    int cls[SP_RPT], rank[SP_RPT];
#pragma unroll
    for (int k = 0; k < SP_RPT; k++) {
        const long long row = r0 + (long long)k * BLK;
        cls[k] = 0;
        rank[k] = 0;
        if (row < m) {
            cls[k] = sp_class(irp[row + 1] - irp[row]);
            rank[k] = atomicAdd(&h[cls[k]], 1);
        }
    }
    __syncthreads();
    if (threadIdx.x < SP_NB && h[threadIdx.x]) base[threadIdx.x] = off[threadIdx.x] + atomicAdd(&cur[threadIdx.x], h[threadIdx.x]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < SP_RPT; k++) {
        const long long row = r0 + (long long)k * BLK;
        if (row < m) perm[base[cls[k]] + rank[k]] = (int)row;
    }
}
}  // namespace mph
