// The rest of the reference's BLAS surface: the v1 vector and matrix routines built
// from the element kernels of src/mpvector.cuh / src/mpmatrix.cuh (multiply by a
// scalar or a vector element, the v1 add and sub, each with its round), the absolute-
// sum and maximum-absolute-value reductions of src/mpreduct.cuh:120-286 and
// src/blas/genorm.cuh, and the v2 kernels (src/blas/v2). Each reference kernel pair
// (esi + digits) and its round kernel run here as one pass per element; the arithmetic
// is the library's lane-group mp_mul, mp_add and v1 add (group.hpp).
#pragma once

#include "kernels.hpp"

namespace mph {

// ---------------------------------------------------------------- v1 element maps
// w(i, j) = x(i, j) * s (V1_MUL), x(i, j) + y(i, j) (V1_ADD, the v1 vector / matrix add
// without the sign-detecting MRC) or x(i, j) - y(i, j) (V1_SUB: y's sign flipped, then the
// same add, src/mpvector.cuh:487-560) for i < m, j < n, where element (i, j) of an SoA
// array is at bidx(i, m, inc) + j * ld (a vector is a single column, n = 1). The factor s
// of V1_MUL is S[0] (SF_SCALAR: vec2scal / mat2scal), S[bidx(i, m, incs)] (SF_ROW:
// mat2vec_left_scal) or S[bidx(j, n, incs)] (SF_COL: mat2vec_right_scal, cvec2rvec).
template <int N, int L, int OP, int SF>
__global__ __launch_bounds__(BLK) void k_v1map(const DevConsts *__restrict__ C, int m, int n, SoA S, int incs, SoA X,
                                               int ldx, int incx, SoA Y, int ldy, int incy, SoA W, int ldw,
                                               int incw) {
    S = rd(S);
    X = rd(X);
    Y = rd(Y);
    W = rd(W);
    const long long e = group_id<L>();
    if (e >= (long long)m * n) return;
    const int i = (int)(e % m), j = (int)(e / m);
    Grp<N, L> g(C);
    MP<N / L> a, b, r;
    load_soa(g, a, X, bidx(i, m, incx) + (long long)j * ldx);
    if constexpr (OP == V1_MUL) {
        const long long is = SF == SF_SCALAR ? 0 : (SF == SF_ROW ? bidx(i, m, incs) : bidx(j, n, incs));
        load_soa(g, b, S, is);
        g.mul(r, a, b);
    } else {
        load_soa(g, b, Y, bidx(i, m, incy) + (long long)j * ldy);
        if (OP == V1_SUB) b.sign ^= 1;
        g.vadd(r, a, b);
    }
    store_soa(g, r, W, bidx(i, m, incw) + (long long)j * ldw);
}

// ---------------------------------------------------------------- maximum absolute value
// mp_array_reduce_max_abs_kernel1/2 (src/mpreduct.cuh:199-286): virtual thread v keeps the
// first of terms v, v + nvt, ... whose magnitude exceeds the kept one (mp_cmp_abs == 1),
// starting from zero; MODE 0 reads an SoA vector with increment incx > 0, MODE 1 AoS
// partials. The halving tree keeps the right partial only when it is larger.
template <int N, int L, int MODE>
__global__ __launch_bounds__(BLK) void k_maxabs_fold(const DevConsts *__restrict__ C, long long cnt, long long nvt,
                                                     SoA xs, int incx, AoSRef<N> xa, AoSRef<N> out) {
    xs = rd(xs);
    const long long v = group_id<L>();
    if (v >= nvt) return;
    Grp<N, L> g(C);
    MP<N / L> acc, p;
    g.zero(acc);
    for (long long i = v; i < cnt; i += nvt) {
        if (MODE == 0)
            load_soa(g, p, xs, i * incx);
        else
            load_aos(g, p, xa, i);
        if (g.template cmp<true>(p, acc) == 1) acc = p;
    }
    store_aos(g, acc, out, v);
}
// the tree over each block of b partials; the block's maximum is stored with sign 0
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_maxabs_tree(const DevConsts *__restrict__ C, AoSRef<N> part, int b, unsigned p2,
                                                     AoSRef<N> out) {
    const long long base = (long long)blockIdx.x * b;
    const int W = BLK / L;
    const int gl = threadIdx.x / L;
    Grp<N, L> g(C);
    for (unsigned s = p2 >> 1; s >= 1; s >>= 1) {
        for (int t = gl; t < (int)s; t += W) {
            if (t + (int)s < b) {
                MP<N / L> a, c;
                load_aos(g, a, part, base + t);
                load_aos(g, c, part, base + t + s);
                if (g.template cmp<true>(c, a) == 1) store_aos(g, c, part, base + t);
            }
        }
        __threadfence_block();
        __syncthreads();
    }
    if (gl == 0) {
        MP<N / L> a;
        load_aos(g, a, part, base);
        a.sign = 0;
        store_aos(g, a, out, blockIdx.x);
    }
}

// ---------------------------------------------------------------- matrix row / column sums of |a|
// matrix_row_sum_abs_kernel (COLS false: one-thread-per-column strides over row o) and
// matrix_col_sum_abs_kernel (COLS true), src/blas/genorm.cuh:38-139: virtual thread
// (o, t) adds |A| entries t, t + b1, ... of row / column o from zero (mp_add_abs: the
// absolute values are nonnegative, so the add is mp_add); part[o * b1 + t]
template <int N, int L, bool COLS>
__global__ __launch_bounds__(BLK) void k_matsum_fold(const DevConsts *__restrict__ C, int m, int n, SoA A, int lda,
                                                     int b1, AoSRef<N> part) {
    A = rd(A);
    const long long v = group_id<L>();
    const int outs = COLS ? n : m, len = COLS ? m : n;
    if (v >= (long long)outs * b1) return;
    const int o = (int)(v / b1), t = (int)(v % b1);
    Grp<N, L> g(C);
    MP<N / L> acc, p;
    g.zero(acc);
    for (int i = t; i < len; i += b1) {
        load_soa(g, p, A, COLS ? (long long)i + (long long)o * lda : (long long)o + (long long)i * lda);
        p.sign = 0;
        g.add(acc, acc, p);
    }
    store_aos(g, acc, part, v);
}

// ---------------------------------------------------------------- v2 element maps
// SCAL: r_i = alpha x_i (src/blas/v2/scal_v2.cuh:37-48); GER: A(i,j) += (alpha x_i) y_j
// (ger_v2.cuh:49-68); SYR: A(i,j) += (x_i x_j) alpha (syr_v2.cuh:58-76); SYR2: A(i,j) += (x_i y_j) alpha
// + (y_i x_j) alpha (syr2_v2.cuh:59-83); SYR and SYR2 only on the upper (row <= col) or lower
// (col <= row) triangle of the n x n matrix. One group per element: the reference's
// kernels do the same work when their launch covers the matrix.
template <int N, int L, int OP>
__global__ __launch_bounds__(BLK) void k_v2map(const DevConsts *__restrict__ C, int m, int n, int upper,
                                               AoSRef<N> alpha, AoSRef<N> x, int incx, AoSRef<N> y, int incy,
                                               AoSRef<N> A, int lda, AoSRef<N> r) {
    const long long e = group_id<L>();
    if (e >= (long long)m * n) return;
    const int row = (int)(e % m), col = (int)(e / m);
    if (OP == V2_SYR || OP == V2_SYR2)
        if (upper ? row > col : col > row) return;
    Grp<N, L> g(C);
    MP<N / L> a, u, w, t1, t2;
    load_aos(g, a, alpha, 0);
    if constexpr (OP == V2_SCAL) {
        load_aos(g, u, x, row);
        g.mul(t1, a, u);
        store_aos(g, t1, r, row);
        return;
    } else {
        const long long ia = (long long)row + (long long)col * lda;
        if constexpr (OP == V2_GER) {
            load_aos(g, u, x, bidx(row, m, incx));
            load_aos(g, w, y, bidx(col, n, incy));
            g.mul(t1, a, u);
            g.mul(t2, t1, w);
        } else if constexpr (OP == V2_SYR) {
            load_aos(g, u, x, bidx(row, n, incx));
            load_aos(g, w, x, bidx(col, n, incx));
            g.mul(t1, u, w);
            g.mul(t2, t1, a);
        } else {
            MP<N / L> p, q;
            load_aos(g, u, x, bidx(row, n, incx));
            load_aos(g, w, y, bidx(col, n, incy));
            g.mul(p, u, w);
            g.mul(t1, p, a);
            load_aos(g, u, y, bidx(row, n, incy));
            load_aos(g, w, x, bidx(col, n, incx));
            g.mul(p, u, w);
            g.mul(q, p, a);
            g.add(t2, t1, q);
        }
        load_aos(g, u, A, ia);
        g.add(t1, u, t2);
        store_aos(g, t1, A, ia);
    }
}

// symv (src/blas/v2/symv_v2.cuh:57-95): y_row = (beta y_row) + alpha (sum over col of
// x_col A(row, col)), A(row, col) read from the stored triangle; one group per row
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_symv_v2(const DevConsts *__restrict__ C, int upper, int n, AoSRef<N> alpha,
                                                 AoSRef<N> A, int lda, AoSRef<N> x, int incx, AoSRef<N> beta,
                                                 AoSRef<N> y, int incy) {
    const long long row = group_id<L>();
    if (row >= n) return;
    Grp<N, L> g(C);
    MP<N / L> s, p, u, w;
    g.zero(s);
    for (int col = 0; col < n; col++) {
        const bool st = (row <= col) == (upper != 0);  // A(row, col) lies in the stored triangle
        const long long ia = st ? row + (long long)col * lda : col + row * (long long)lda;
        load_aos(g, u, x, bidx(col, n, incx));
        load_aos(g, w, A, ia);
        g.mul(p, u, w);
        g.add(s, s, p);
    }
    const long long iy = bidx(row, n, incy);
    load_aos(g, u, beta, 0);
    load_aos(g, w, y, iy);
    g.mul(p, u, w);  // beta * y
    load_aos(g, u, alpha, 0);
    g.mul(w, u, s);  // alpha * sums
    g.add(s, p, w);
    store_aos(g, s, y, iy);
}

// syrk (src/blas/v2/syrk_v2.cuh:61-85): C(row, col) = (sum over i of op(A)(row, i) op(A)(col, i))
// alpha + C(row, col) beta on the upper or lower triangle; one group per element
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_syrk_v2(const DevConsts *__restrict__ C, int upper, int trans, int n, int k,
                                                 AoSRef<N> alpha, AoSRef<N> A, int lda, AoSRef<N> beta, AoSRef<N> Cm,
                                                 int ldc) {
    const long long e = group_id<L>();
    if (e >= (long long)n * n) return;
    const int row = (int)(e % n), col = (int)(e / n);
    if (upper ? row > col : col > row) return;
    Grp<N, L> g(C);
    MP<N / L> s, p, u, w;
    g.zero(s);
    for (int i = 0; i < k; i++) {
        const long long ia = trans ? i + (long long)lda * row : row + (long long)lda * i;
        const long long ib = trans ? i + (long long)lda * col : col + (long long)lda * i;
        load_aos(g, u, A, ia);
        load_aos(g, w, A, ib);
        g.mul(p, u, w);
        g.add(s, s, p);
    }
    load_aos(g, u, alpha, 0);
    g.mul(p, s, u);  // sum * alpha
    const long long ic = row + (long long)col * ldc;
    load_aos(g, u, Cm, ic);
    load_aos(g, w, beta, 0);
    g.mul(s, u, w);  // C * beta
    g.add(u, p, s);
    store_aos(g, u, Cm, ic);
}

}  // namespace mph
