// Precision form of gemm on INT8 MFMA (mpres_gemm_mfma).
//
// C = alpha * op(A) op(B) + beta * C, with the product formed as one exact sum
// per entry instead of the reference's rounded step-by-step chain:
//   1. every row of op(A) (column of op(B)) gets a common exponent E = the
//      row's minimum exponent + a shift s; s = 0 unless the aligned
//      magnitudes would exceed the bit budget that keeps |sum| < M/4, in
//      which case the row is truncated by s bits (the tile alignment rounding);
//   2. the aligned residues x_r 2^(e - E) mod m_r (signed, mod m_r) are split
//      into 7-bit limbs and stored as int8 planes;
//   3. for each modulus the limb products run on v_mfma_i32_32x32x32_i8 with
//      INT32 accumulation, grouped by limb shift a + b, and are recombined and
//      reduced mod m_r: the residues of the exact sum;
//   4. per entry the interval evaluation of the sum is computed from its
//      residues, the sign read off it (|sum| < M/4), the sum rounded once to
//      the working precision, and the epilogue (sum alpha) + (C beta) applied
//      with the reference's mp_mul / mp_add.
// Residues stay exact; only the truncation of step 1 (when needed) and the
// final rounding move the result away from the exact value.
#pragma once

#include "kernels.hpp"

namespace mph {

typedef int mf_v4i __attribute__((ext_vector_type(4)));
typedef int mf_v16i __attribute__((ext_vector_type(16)));

constexpr int MF_T = 32;     // MFMA tile (rows, columns and k per instruction)
constexpr int MF_LB = 7;     // bits per limb (limbs are 0..127, valid int8)

// bit length bound of an element from its interval upper bound: X < 2^(exp+1) M < 2^(exp+2+log2m)
template <int N, class VM>
__device__ __forceinline__ int mf_bits(const VM &X, long long i, int log2m) {
    int s, e;
    Er lo, up;
    scalars_of<N>(X, i, s, e, lo, up);
    return up.frac != 0 ? up.exp + 2 + log2m : -0x40000000;
}

// maximum aligned bit length of each row of op(A) (SIDE 0) or column of op(B) (SIDE 1):
// max over i of bits(x) + e(x) - emin[j], combined by atomic maximum into out
template <int N, int L, class VM, int SIDE>
__global__ __launch_bounds__(BLK) void k_mf_bits(const DevConsts *__restrict__ C, int cnt, int k, int t, VM X, int ld,
                                                 const int *__restrict__ emin, int *out) {
    if constexpr (std::is_same<VM, SoA>::value) X = rd(X);
    constexpr int S = BLK / EMIN_J;
    const long long j = (long long)blockIdx.x * EMIN_J + threadIdx.x % EMIN_J;
    const int i0 = blockIdx.y * EMIN_K, i1 = (i0 + EMIN_K) < k ? i0 + EMIN_K : k;
    if (j >= cnt) return;
    int b = -0x40000000;
    for (int i = i0 + (int)threadIdx.x / EMIN_J; i < i1; i += S) {
        const long long idx = op_index<SIDE>(t, (int)j, i, ld);
        const int v = mf_bits<N>(X, idx, C->log2m) + exp_of<N>(X, idx) - emin[j];
        b = v > b ? v : b;
    }
    if (b != -0x40000000) atomicMax(out + j, b);
}

// shift of each row / column: s = max(0, bits - half); emin becomes the effective
// exponent E + s (0 for a row or column of zeros), bits the aligned bound min(bits, half)
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_mf_shift(int cnt, int half, int *emin_eff, int *bits) {
    const long long j = (long long)blockIdx.x * BLK + threadIdx.x;
    if (j >= cnt) return;
    const int b = bits[j];
    const int s = b > half ? b - half : 0;
    emin_eff[j] = emin_eff[j] == 0x7fffffff ? 0 : emin_eff[j] + s;
    bits[j] = b < half ? b : half;
}

// aligned, signed residues of op(A) rows (SIDE 0) or op(B) columns (SIDE 1) as int8 limb
// planes P[r][limb][j][kp] (kp = 0 .. KP-1, zero padded; j = 0 .. JP-1, zero padded)
template <int N, int L, class VM, int SIDE>
__global__ __launch_bounds__(BLK) void k_mf_prescale(const DevConsts *__restrict__ C, int cnt, int JP, int k, int KP,
                                                     int nl, int t, VM X, int ld, const int *__restrict__ eff,
                                                     signed char *__restrict__ P) {
    if constexpr (std::is_same<VM, SoA>::value) X = rd(X);
    const long long g = (long long)blockIdx.x * BLK + threadIdx.x;
    if (g >= (long long)JP * KP) return;
    const int kp = (int)(g % KP), j = (int)(g / KP);
    const size_t plane = (size_t)JP * KP;
    if (j >= cnt || kp >= k) {
        for (int r = 0; r < N; r++)
            for (int a = 0; a < nl; a++) P[((size_t)r * nl + a) * plane + g] = 0;
        return;
    }
    const long long idx = op_index<SIDE>(t, j, kp, ld);
    const int d = exp_of<N>(X, idx) - eff[j];
    const int s = sign_at<N>(X, idx);
    const int *dig = digits_of<N>(X, idx);
    int x[N];
    for (int r = 0; r < N; r++) x[r] = dig[r];
    if (d < 0) {
        // truncation of the row's alignment: floor(X / 2^-d)
        Core<N, 1> core{C, C->pow2, 0};
        rns_scale2pow<N, 1, false>(core, x, (unsigned)(-d));
    }
    for (int r = 0; r < N; r++) {
        const int m = C->m[r];
        int v = d > 0 ? mulmod(x[r], pow2mod(C, d, r), C->md[r], C->minv[r]) : x[r];
        v = (s && v) ? m - v : v;
        for (int a = 0; a < nl; a++) P[((size_t)r * nl + a) * plane + g] = (signed char)((v >> (MF_LB * a)) & 127);
    }
}

// Limb products per modulus on MFMA. Workgroup: 4 waves, a 64 x 64 block of C
// (each wave a 32 x 32 tile); blockIdx.z = modulus. Operands are swapped
// (B^T as the first, A^T as the second) so that a lane holds one row of C.
// Output: S[r][e], e = row + m * col, the residue of the exact sum mod m_r.
template <int N, int L, int NL>
__global__ __launch_bounds__(BLK) void k_mf_gemm(const DevConsts *__restrict__ C, int m, int n, int MP, int NP, int KP,
                                                 const signed char *__restrict__ PA, const signed char *__restrict__ PB,
                                                 int *__restrict__ S) {
    constexpr int NG = 2 * NL - 1;  // limb shift groups
    const int r = blockIdx.z;
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    const int row0 = blockIdx.y * 64 + 32 * (w & 1), col0 = blockIdx.x * 64 + 32 * (w >> 1);
    const size_t pa = (size_t)MP * KP, pb = (size_t)NP * KP;
    const signed char *A = PA + (size_t)r * NL * pa + (size_t)(row0 + l % 32) * KP + 16 * (l / 32);
    const signed char *B = PB + (size_t)r * NL * pb + (size_t)(col0 + l % 32) * KP + 16 * (l / 32);
    mf_v16i acc[NG];
#pragma unroll
    for (int q = 0; q < NG; q++) acc[q] = mf_v16i{};
    if (row0 < MP && col0 < NP) {
        for (int k0 = 0; k0 < KP; k0 += MF_T) {
            mf_v4i fa[NL], fb[NL];
#pragma unroll
            for (int a = 0; a < NL; a++) {
                fa[a] = *(const mf_v4i *)(A + a * pa + k0);
                fb[a] = *(const mf_v4i *)(B + a * pb + k0);
            }
#pragma unroll
            for (int a = 0; a < NL; a++)
#pragma unroll
                for (int b = 0; b < NL; b++)
                    acc[a + b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fb[b], fa[a], acc[a + b], 0, 0, 0);
        }
    }
    const double md = C->md[r], mv = C->minv[r];
    const int row = row0 + l % 32;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int col = col0 + 8 * (q / 4) + 4 * (l / 32) + q % 4;
        double v = 0.0;
#pragma unroll
        for (int g = NG - 1; g >= 0; g--) v = (double)modd(v * (double)(1 << MF_LB) + (double)acc[g][q], md, mv);
        if (row < m && col < n) S[(size_t)r * m * n + (size_t)row + (size_t)m * col] = (int)v;
    }
}

// per entry: the exact sum from its residues, rounded once, then C = (sum alpha) + (C beta)
// The interval evaluation of the sum is the reference's full evaluation of the
// residues of |sum| 2^K0, exponents lowered by K0: with the bit bounds of the
// entry's row and column, |sum| 2^K0 < M/2, so the evaluation starts close to
// the top of the range instead of refining up from |sum| / M.
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_mf_finish(const DevConsts *__restrict__ C_, int m, int n, int lk,
                                                   const int *__restrict__ S, const int *__restrict__ EA,
                                                   const int *__restrict__ EB, const int *__restrict__ BA,
                                                   const int *__restrict__ BB, SoA alpha, SoA beta, SoA Cm, int ldc) {
    alpha = rd(alpha);
    beta = rd(beta);
    Cm = rd(Cm);
    constexpr int R = N / L;
    const long long e = group_id<L>();
    if (e >= (long long)m * n) return;
    const int row = (int)(e % m), col = (int)(e / m);
    Grp<N, L> g(C_);
    MP<R> sum, a, b, cv, s2, c2, res;
    const int log2m = C_->log2m;
    int k0 = log2m - 1 - (BA[row] + BB[col] + lk);
    k0 = k0 < 0 ? 0 : (k0 > log2m - 1 ? log2m - 1 : k0);
    const int *p2 = C_->pow2 + (size_t)k0 * N;
    int sc[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
        sum.d[q] = S[(size_t)g.ix(q) * m * n + e];
        sc[q] = mulmod(sum.d[q], p2[g.ix(q)], g.md(q), g.mv(q));
    }
    sum.exp = EA[row] + EB[col];
    Er lo, up;
    g.eval_full(sc, lo, up);
    sum.sign = 0;
    if (up.frac != 0 && up.exp >= -1) {
        // 2^K0 X / M >= 1/2: the residues hold M - |sum| (|sum| 2^K0 < M/2)
#pragma unroll
        for (int q = 0; q < R; q++) {
            sum.d[q] = sum.d[q] ? g.mi(q) - sum.d[q] : 0;
            sc[q] = sc[q] ? g.mi(q) - sc[q] : 0;
        }
        sum.sign = 1;
        g.eval_full(sc, lo, up);
    }
    if (lo.frac != 0) lo.exp -= k0;
    if (up.frac != 0) up.exp -= k0;
    sum.ev0 = lo;
    sum.ev1 = up;
    g.round_if_needed(sum);
    const long long ic = (long long)col * ldc + row;
    load_soa(g, a, alpha, 0);
    load_soa(g, b, beta, 0);
    load_soa(g, cv, Cm, ic);
    g.mul(s2, sum, a);
    g.mul(c2, cv, b);
    g.add(res, s2, c2);
    store_soa(g, res, Cm, ic);
}

}  // namespace mph
