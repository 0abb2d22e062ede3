// Batched form (mpres_batch_run): many independent dot and gemv problems in one
// persistent launch.
//
// The host splits the batch into work items (a whole dot, or a run of output
// rows of a gemv) and orders them by descending work. Each workgroup takes the
// next item from a device counter, runs it to the end with workgroup barriers
// between its phases, and takes another, until the counter passes the item
// count; every workgroup therefore exits. Items share no data (a gemv's rows
// are split so that each item owns its outputs and recomputes alpha * x for
// itself), so there are no waits between workgroups.
//
// Each item performs the same mp operations in the same order as the single
// call it stands for, so its results are bit-identical to it:
//   dot  (mpres_dot, src/blas/dot.cuh:84-107 with src/mpreduct.cuh:38-112):
//        grid3 * block3 strided product folds, a halving tree per block of
//        block3, block3 strided folds of the grid3 block results, a last tree;
//   gemv (mpres_gemv, src/blas/gemv.cuh:150-268): bx = x * alpha, y = y * beta,
//        block3 strided folds of A-entries times bx per output, a halving tree
//        per output, y = y + sum.
// The chains run the full arithmetic (every rounding inline), so there is no
// separate fixup pass. Partial results live in the item's scratch (global
// memory, cached), as in the single-call kernels.
#pragma once

#include "kernels.hpp"

namespace mph {

__device__ __forceinline__ void wg_sync() {
    __threadfence_block();
    __syncthreads();
}

// k_tree's halving tree over nb blocks of b partials each (block q at part[q * b]), all
// blocks level by level: for s = nextPow2(b)/2 .. 1: p[t] = p[t] + p[t + s] for t < s, t + s < b
template <int N, int L>
__device__ void batch_tree(const Grp<N, L> &g, AoSRef<N> part, long long nb, int b) {
    constexpr int W = BLK / L;
    const int gl = threadIdx.x / L;
    unsigned p2 = 1;
    while (p2 < (unsigned)b) p2 <<= 1;
    for (unsigned s = p2 >> 1; s >= 1; s >>= 1) {
        const long long tot = nb * s;
        for (long long q = gl; q < tot; q += W) {
            const long long blk = q / s;
            const int t = (int)(q % s);
            if (t + (int)s < b) {
                MP<N / L> a, c, r;
                load_aos(g, a, part, blk * b + t);
                load_aos(g, c, part, blk * b + t + s);
                g.add(r, a, c);
                store_aos(g, r, part, blk * b + t);
            }
        }
        wg_sync();
    }
}

// dot: scratch = grid3 * block3 partials, then block3 second-pass partials
template <int N, int L>
__device__ void batch_dot(const Grp<N, L> &g, const BatchItem &it, char *ws) {
    constexpr int W = BLK / L;
    const int gl = threadIdx.x / L;
    const SoA x = rd(it.x), y = rd(it.y);
    const int n = it.n, g3 = it.g3, b3 = it.b3;
    const long long G = (long long)g3 * b3;
    const AoSRef<N> part{ws}, part2{ws + G * AoSRef<N>::STRIDE};
    for (long long v = gl; v < G; v += W) {
        MP<N / L> acc;
        g.zero(acc);
        for (long long i = v; i < n; i += G)
            g.template fma_step<true>(acc, x, bidx(i, n, it.incx), y, bidx(i, n, it.incy));
        store_aos(g, acc, part, v);
    }
    wg_sync();
    batch_tree<N, L>(g, part, g3, b3);
    for (int t = gl; t < b3; t += W) {
        MP<N / L> acc, p;
        g.zero(acc);
        for (int j = t; j < g3; j += b3) {
            load_aos(g, p, part, (long long)j * b3);
            g.add(acc, acc, p);
        }
        store_aos(g, acc, part2, t);
    }
    wg_sync();
    batch_tree<N, L>(g, part2, 1, b3);
    if (gl == 0) {
        MP<N / L> a;
        load_aos(g, a, part2, 0);
        const SoA r = rd(it.r);
        store_soa(g, a, r, 0);
    }
}

// gemv rows [r0, r1): scratch = bx (lx numbers), then (r1 - r0) * block3 partials
template <int N, int L, bool TRANS>
__device__ void batch_gemv(const Grp<N, L> &g, const BatchItem &it, char *ws) {
    constexpr int W = BLK / L;
    const int gl = threadIdx.x / L;
    const SoA alpha = rd(it.alpha), A = rd(it.A), x = rd(it.x), beta = rd(it.beta), y = rd(it.y);
    const int lx = TRANS ? it.m : it.n, ly = TRANS ? it.n : it.m;
    const int r0 = it.r0, h = it.r1 - it.r0, b3 = it.b3;
    const long long lda = it.lda;
    const AoSRef<N> bx{ws}, part{ws + (long long)lx * AoSRef<N>::STRIDE};
    {
        MP<N / L> a, v, r;
        load_soa(g, a, alpha, 0);
        for (int k = gl; k < lx; k += W) {
            load_soa(g, v, x, bidx(k, lx, it.incx));
            g.mul(r, v, a);
            store_aos(g, r, bx, k);
        }
        load_soa(g, a, beta, 0);
        for (int o = r0 + gl; o < r0 + h; o += W) {
            const long long iy = bidx(o, ly, it.incy);
            load_soa(g, v, y, iy);
            g.mul(r, v, a);
            store_soa(g, r, y, iy);
        }
    }
    wg_sync();
    // without transpose consecutive groups take consecutive rows (one column run of A per step)
    const long long nv = (long long)h * b3;
    for (long long q = gl; q < nv; q += W) {
        const int ol = TRANS ? (int)(q / b3) : (int)(q % h);
        const int t = TRANS ? (int)(q % b3) : (int)(q / h);
        const long long o = r0 + ol;
        MP<N / L> acc;
        g.zero(acc);
        for (int j = t; j < lx; j += b3) g.template fma_step<true>(acc, A, TRANS ? (o * lda + j) : (j * lda + o), bx, j);
        store_aos(g, acc, part, (long long)ol * b3 + t);
    }
    wg_sync();
    batch_tree<N, L>(g, part, h, b3);
    for (int ol = gl; ol < h; ol += W) {
        const long long iy = bidx(r0 + ol, ly, it.incy);
        MP<N / L> a, s, r;
        load_soa(g, a, y, iy);
        load_aos(g, s, part, (long long)ol * b3);
        g.add(r, a, s);
        store_soa(g, r, y, iy);
    }
}

// the persistent kernel: *queue must be 0 at launch (the launcher clears it first)
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_batch(const DevConsts *__restrict__ C, const BatchItem *__restrict__ items,
                                               int count, int *queue, char *ws) {
    __shared__ int next;
    Grp<N, L> g(C);
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(queue, 1);
        __syncthreads();
        const int i = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();
        if (i >= count) return;
        const BatchItem &it = items[i];
        char *w = ws + it.ws;
        if (it.kind == BATCH_DOT)
            batch_dot<N, L>(g, it, w);
        else if (it.kind == BATCH_GEMV_N)
            batch_gemv<N, L, false>(g, it, w);
        else
            batch_gemv<N, L, true>(g, it, w);
        wg_sync();
    }
}

}  // namespace mph
