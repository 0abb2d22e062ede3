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

constexpr int MF_T = LIMB_T;     // MFMA tile (rows, columns and k per instruction)
constexpr int MF_LB = LIMB_BITS; // bits per limb (limbs are 0..127, valid int8)
constexpr int MF_KCH = 8192;     // k per INT32 accumulation chunk: NL 127^2 MF_KCH < 2^31 for NL <= 5
constexpr int MF_KONE = 16384;   // k up to which one INT32 accumulation suffices (NL <= 5)
constexpr int MF_NONE = -0x40000000;  // no element (bit bounds of empty rows, columns, bands)

// bit length bound of an element from its interval upper bound: X < 2^(exp+1) M < 2^(exp+2+log2m)
template <int N, class VM>
__device__ __forceinline__ int mf_bits(const VM &X, long long i, int log2m) {
    int s, e;
    Er lo, up;
    scalars_of<N>(X, i, s, e, lo, up);
    return up.frac != 0 ? up.exp + 2 + log2m : MF_NONE;
}

// Highest bit position (bits + exponent) of each row of op(A) (SIDE 0) or column of op(B)
// (SIDE 1), over its nonzero elements (LOW: only those with exponent below eh[j]),
// combined by atomic maximum into out (filled with MF_NONE first).
template <int N, int L, class VM, int SIDE, bool LOW>
__global__ __launch_bounds__(BLK) void k_mf_top(const DevConsts *__restrict__ C, int cnt, int k, int t, VM X, int ld,
                                                const int *__restrict__ eh, int *out) {
    if constexpr (std::is_same<VM, SoA>::value) X = rd(X);
    constexpr int S = BLK / EMIN_J;
    const long long j = (long long)blockIdx.x * EMIN_J + threadIdx.x % EMIN_J;
    const int i0 = blockIdx.y * EMIN_K, i1 = (i0 + EMIN_K) < k ? i0 + EMIN_K : k;
    if (j >= cnt) return;
    int b = MF_NONE;
    for (int i = i0 + (int)threadIdx.x / EMIN_J; i < i1; i += S) {
        const long long idx = op_index<SIDE>(t, (int)j, i, ld);
        const int bits = mf_bits<N>(X, idx, C->log2m);
        if (bits == MF_NONE) continue;
        const int e = exp_of<N>(X, idx);
        if (LOW && e >= eh[j]) continue;
        const int v = bits + e;
        b = v > b ? v : b;
    }
    if (b != MF_NONE) atomicMax(out + j, b);
}

// Exponent bands of each row / column. With half the bit budget of one side:
//   high band: exponent eh = max(emin, top - half), aligned bits bh = top - eh <= half; it
//              holds the elements with exponent >= eh (exactly: no truncation);
//   low band:  the elements with exponent < eh; exponent el = emin + max(0, topl - emin - half),
//              aligned bits bl = min(topl - emin, half); elements below el are truncated.
// Both are e = emin + max(0, t - emin - half), b = min(t - emin, half) for t = top (high band)
// or t = topl (low band). No element (t = MF_NONE, or a row of zeros): e = 0, b = MF_NONE.
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_mf_band(int cnt, int half, const int *__restrict__ emin,
                                                 const int *__restrict__ top, int *e_out, int *b_out) {
    const long long j = (long long)blockIdx.x * BLK + threadIdx.x;
    if (j >= cnt) return;
    const int lo = emin[j], tp = top[j];
    if (lo == 0x7fffffff || tp == MF_NONE) {
        e_out[j] = 0;
        b_out[j] = MF_NONE;
        return;
    }
    const int span = tp - lo;
    e_out[j] = lo + (span > half ? span - half : 0);
    b_out[j] = span < half ? span : half;
}

// aligned, signed residues of op(A) rows (SIDE 0) or op(B) columns (SIDE 1) as int8 limb
// planes P[band][r][limb] (band 1 high, 0 low; laid out as the planes of k_gemm_limbs, zero
// padded); an element goes to one band, the other band's planes hold 0 there. A row or
// column without a low band (bl[j] == MF_NONE) leaves its low-band planes unwritten: no
// entry of it reads a low-band product (the finish takes only the band pairs present),
// and its padding beyond cnt is never read either.
// At least four waves per SIMD: left alone the compiler keeps all 32 residues and their
// loads in flight at 170 registers (two waves), too few to hide the memory latency.
template <int N, int L, class VM, int SIDE>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(4))) void
k_mf_prescale(const DevConsts *__restrict__ C, int cnt, int JP, int k, int KP, int nl, int t, VM X, int ld,
              const int *__restrict__ eh, const int *__restrict__ el, const int *__restrict__ bl,
              signed char *__restrict__ P) {
    if constexpr (std::is_same<VM, SoA>::value) X = rd(X);
    const long long g = (long long)blockIdx.x * BLK + threadIdx.x;
    if (g >= (long long)JP * KP) return;
    int j, kp;
    limb_elem(g, JP, j, kp);
    const size_t plane = (size_t)JP * KP, band = (size_t)N * nl * plane;
    signed char *PL = P, *PH = P + band;
    const bool low = j < cnt && bl[j] != MF_NONE;  // this row / column has a low band
    if (j >= cnt || kp >= k) {
        for (int r = 0; r < N; r++)
            for (int a = 0; a < nl; a++) {
                if (low) PL[((size_t)r * nl + a) * plane + g] = 0;
                PH[((size_t)r * nl + a) * plane + g] = 0;
            }
        return;
    }
    const long long idx = op_index<SIDE>(t, j, kp, ld);
    int s, e;
    Er lo, up;
    scalars_of<N>(X, idx, s, e, lo, up);
    const bool high = e >= eh[j];
    signed char *Pw = high ? PH : PL, *Pz = high ? PL : PH;
    const bool wz = !high || low;  // the other band's plane needs its zero
    const int d = e - (high ? eh[j] : el[j]);
    // the element's residues at once (SoA records are 16-byte aligned)
    int x[N];
    if constexpr (std::is_same<VM, SoA>::value) {
        const int4 *p4 = (const int4 *)digits_of<N>(X, idx);
#pragma unroll
        for (int q = 0; q < N / 4; q++) {
            const int4 v4 = p4[q];
            x[4 * q] = v4.x;
            x[4 * q + 1] = v4.y;
            x[4 * q + 2] = v4.z;
            x[4 * q + 3] = v4.w;
        }
    } else {
        const int *dig = digits_of<N>(X, idx);
        for (int r = 0; r < N; r++) x[r] = dig[r];
    }
    if (up.frac == 0 || -d > C->log2m + 1) {
        // zero, or truncated to zero (X < M < 2^(log2m + 1))
        for (int r = 0; r < N; r++) x[r] = 0;
    } else if (d < 0) {
        // truncation of the low band's alignment: floor(X / 2^-d)
        Core<N, 1> core{C, C->pow2, 0};
        rns_scale2pow<N, 1, false>(core, x, (unsigned)(-d));
    }
    for (int r = 0; r < N; r++) {
        const int m = C->m[r];
        int v = (d > 0 && x[r]) ? mulmod(x[r], pow2mod(C, d, r), C->md[r], C->minv[r]) : x[r];
        v = (s && v) ? m - v : v;
        for (int a = 0; a < nl; a++) {
            Pw[((size_t)r * nl + a) * plane + g] = (signed char)((v >> (MF_LB * a)) & 127);
            if (wz) Pz[((size_t)r * nl + a) * plane + g] = 0;
        }
    }
}

// Limb products per modulus on MFMA. Workgroup: 4 waves, a 64 x 64 block of C
// (each wave a 32 x 32 tile); blockIdx.z = modulus. Operands are swapped
// (B^T as the first, A^T as the second) so that a lane holds one row of C.
// Per k-step of 32 the workgroup's 64-row panel of A and 64-column panel of B (all
// limbs) pass through LDS once, double-buffered: the next step's pieces are loaded
// into registers while this step's products run, then stored to the other buffer.
// Each wave reads its fragments from LDS (consecutive lanes, consecutive 16 bytes).
// Output: S[r][e], e = row + m * col, the residue of the exact sum mod m_r.
// fa / fb (low-band products only): the bit bounds of the rows / columns' low bands; a
// workgroup whose 64 rows (or columns) all lack a low band has nothing to add, and the
// finish does not read its entries, so it returns at once.
// CHUNKED (k > MF_KONE only): the INT32 accumulators are folded into residues every MF_KCH
// k; that needs registers for the running residues, so shorter k takes the single pass.
// The grid covers MP x NP exactly in 64-blocks (MP, NP multiples of 64, KP of 32).
template <int N, int L, int NL, bool CHUNKED>
__global__ __launch_bounds__(BLK) void k_mf_gemm(const DevConsts *__restrict__ C, int m, int n, int MP, int NP, int KP,
                                                 const signed char *__restrict__ PA, const signed char *__restrict__ PB,
                                                 const int *__restrict__ fa, const int *__restrict__ fb,
                                                 int *__restrict__ S) {
    constexpr int NG = 2 * NL - 1;               // limb shift groups
    constexpr int ITEMS = NL * 128;              // 16-byte pieces of one 64 x 32 panel, all limbs
    constexpr int PER = (ITEMS + BLK - 1) / BLK; // pieces per thread
    __shared__ mf_v4i sa[2][ITEMS], sb[2][ITEMS];  // [buffer][(limb * 64 + row) * 2 + half]
    if (fa || fb) {
        __shared__ int live[2];
        if (threadIdx.x < 2) live[threadIdx.x] = 0;
        __syncthreads();
        if (threadIdx.x < 64) {
            const int row = blockIdx.y * 64 + threadIdx.x, col = blockIdx.x * 64 + threadIdx.x;
            if (!fa || (row < m && fa[row] != MF_NONE)) live[0] = 1;
            if (!fb || (col < n && fb[col] != MF_NONE)) live[1] = 1;
        }
        __syncthreads();
        if (!(live[0] && live[1])) return;
    }
    const int r = blockIdx.z;
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    const int rb = blockIdx.y * 64, cb = blockIdx.x * 64;
    const int row0 = rb + 32 * (w & 1), col0 = cb + 32 * (w >> 1);
    const size_t pa = (size_t)MP * KP, pb = (size_t)NP * KP;
    // planes are k-step major (k_gemm_limbs): step k0's 64-row panel of one limb is the
    // 64 * MF_T contiguous bytes at k0 * MP + rb * MF_T
    const signed char *Ab = PA + (size_t)r * NL * pa + (size_t)rb * MF_T;
    const signed char *Bb = PB + (size_t)r * NL * pb + (size_t)cb * MF_T;
    mf_v4i ga[PER], gb[PER];
    // piece q: limb q >> 7, row (column) (q >> 1) & 63, half q & 1 of the k-step's 32 bytes
    auto gload = [&](int k0) {
#pragma unroll
        for (int p = 0; p < PER; p++) {
            const int q = threadIdx.x + p * BLK;
            if (q < ITEMS) {
                const int o = ((q >> 1) & 63) * MF_T + 16 * (q & 1);
                ga[p] = *(const mf_v4i *)(Ab + (size_t)(q >> 7) * pa + (size_t)k0 * MP + o);
                gb[p] = *(const mf_v4i *)(Bb + (size_t)(q >> 7) * pb + (size_t)k0 * NP + o);
            }
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int p = 0; p < PER; p++) {
            const int q = threadIdx.x + p * BLK;
            if (q < ITEMS) {
                sa[buf][q] = ga[p];
                sb[buf][q] = gb[p];
            }
        }
    };
    // this lane's fragment pieces: A row 32 (w & 1) + l % 32, B column 32 (w >> 1) + l % 32, half l / 32
    const int ia = (32 * (w & 1) + l % 32) * 2 + l / 32, ib = (32 * (w >> 1) + l % 32) * 2 + l / 32;
    const double md = C->md[r], mv = C->minv[r];
    mf_v16i acc[NG];
#pragma unroll
    for (int q = 0; q < NG; q++) acc[q] = mf_v16i{};
    double res[CHUNKED ? 16 : 1];  // residues of the chunks done so far
#pragma unroll
    for (int q = 0; q < (CHUNKED ? 16 : 1); q++) res[q] = 0.0;
    gload(0);
    sstore(0);
    __syncthreads();
    int buf = 0;
    for (int k0 = 0; k0 < KP; k0 += MF_T) {
        const bool more = k0 + MF_T < KP;
        if (more) gload(k0 + MF_T);
        mf_v4i xa[NL], xb[NL];
#pragma unroll
        for (int a = 0; a < NL; a++) {
            xa[a] = sa[buf][a * 128 + ia];
            xb[a] = sb[buf][a * 128 + ib];
        }
#pragma unroll
        for (int a = 0; a < NL; a++)
#pragma unroll
            for (int b = 0; b < NL; b++)
                acc[a + b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(xb[b], xa[a], acc[a + b], 0, 0, 0);
        if constexpr (CHUNKED) {
            if ((k0 + MF_T) % MF_KCH == 0 || !more) {
                // the chunk's limb groups recombined (Horner, exact below 2^53) and added mod m
#pragma unroll
                for (int q = 0; q < 16; q++) {
                    double v = 0.0;
#pragma unroll
                    for (int g = NG - 1; g >= 0; g--)
                        v = (double)modd(v * (double)(1 << MF_LB) + (double)acc[g][q], md, mv);
                    v += res[q];
                    res[q] = v >= md ? v - md : v;
                }
#pragma unroll
                for (int q = 0; q < NG; q++) acc[q] = mf_v16i{};
            }
        }
        if (more) sstore(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    const int row = row0 + l % 32;
#pragma unroll
    for (int q = 0; q < 16; q++) {
        const int col = col0 + 8 * (q / 4) + 4 * (l / 32) + q % 4;
        double v;
        if constexpr (CHUNKED) {
            v = res[q];
        } else {
            v = 0.0;
#pragma unroll
            for (int g = NG - 1; g >= 0; g--) v = (double)modd(v * (double)(1 << MF_LB) + (double)acc[g][q], md, mv);
        }
        if (row < m && col < n) S[(size_t)r * m * n + (size_t)row + (size_t)m * col] = (int)v;
    }
}

// Band pair p = 2 bA + bB (band 1 high, 0 low) of an entry: the exact sum S_p of the
// products of the row's band bA with the column's band bB, as an mp number rounded once.
// The interval evaluation is the reference's full evaluation of the residues of
// |S_p| 2^K0, exponents lowered by K0: with the bit bounds ba, bb of the two bands,
// |S_p| 2^K0 < M/2, so the evaluation starts close to the top of the range instead of
// refining up from |S_p| / M. Values above M/2 are negative sums (|S_p| < M/4).
template <int N, int L>
__device__ __forceinline__ void mf_pair_value(const Grp<N, L> &g, MP<N / L> &v, const int *__restrict__ Sp,
                                              long long e, long long mn, int ea, int eb, int ba, int bb, int lk) {
    constexpr int R = N / L;
    const DevConsts *C = g.k.C;
    const int log2m = C->log2m;
    ba = ba == MF_NONE ? 0 : ba;
    bb = bb == MF_NONE ? 0 : bb;
    int k0 = log2m - 1 - (ba + bb + lk);
    k0 = k0 < 0 ? 0 : (k0 > log2m - 1 ? log2m - 1 : k0);
    const int *p2 = C->pow2 + (size_t)k0 * N;
    int sc[R];
#pragma unroll
    for (int q = 0; q < R; q++) {
        v.d[q] = Sp[(size_t)g.ix(q) * mn + e];
        sc[q] = mulmod(v.d[q], p2[g.ix(q)], g.md(q), g.mv(q));
    }
    v.exp = ea + eb;
    Er lo, up;
    g.eval_full(sc, lo, up);
    v.sign = 0;
    if (up.frac != 0 && up.exp >= -1) {
#pragma unroll
        for (int q = 0; q < R; q++) {
            v.d[q] = v.d[q] ? g.mi(q) - v.d[q] : 0;
            sc[q] = sc[q] ? g.mi(q) - sc[q] : 0;
        }
        v.sign = 1;
        g.eval_full(sc, lo, up);
    }
    if (lo.frac != 0) lo.exp -= k0;
    if (up.frac != 0) up.exp -= k0;
    v.ev0 = lo;
    v.ev1 = up;
    g.round_if_needed(v);
}

// Per entry: the band pairs present (both bands non-empty; the high pair always) in the
// order low x low, low x high, high x low, high x high, each rounded once and added
// with mp_add; then C = (sum alpha) + (C beta) with mp_mul / mp_add. With one band per
// row and column (no low bands) this is the single exact sum rounded once.
// Band arrays: E* exponents, B* bit bounds, *H high and *L low band, *A rows, *B columns.
template <int N, int L>
__device__ __forceinline__ void mf_finish_rounded(const Grp<N, L> &g, long long e, long long mn, int row, int col,
                                                  int lk, const int *__restrict__ S, const int *__restrict__ EHA,
                                                  const int *__restrict__ BHA, const int *__restrict__ ELA,
                                                  const int *__restrict__ BLA, const int *__restrict__ EHB,
                                                  const int *__restrict__ BHB, const int *__restrict__ ELB,
                                                  const int *__restrict__ BLB, const SoA &alpha, const SoA &beta,
                                                  const SoA &Cm, int ldc) {
    constexpr int R = N / L;
    MP<R> sum, v, a, b, cv, s2, c2, res;
    const bool la = BLA[row] != MF_NONE, lb = BLB[col] != MF_NONE;
    bool first = true;
#pragma unroll 1
    for (int p = 0; p < 4; p++) {
        const int ba_ = p >> 1, bb_ = p & 1;
        if ((!ba_ && !la) || (!bb_ && !lb)) continue;
        mf_pair_value<N, L>(g, v, S + (size_t)p * N * mn, e, mn, ba_ ? EHA[row] : ELA[row], bb_ ? EHB[col] : ELB[col],
                            ba_ ? BHA[row] : BLA[row], bb_ ? BHB[col] : BLB[col], lk);
        if (first) {
            sum = v;
            first = false;
        } else {
            g.add(sum, sum, v);
        }
    }
    const long long ic = (long long)col * ldc + row;
    load_soa(g, a, alpha, 0);
    load_soa(g, b, beta, 0);
    load_soa(g, cv, Cm, ic);
    g.mul(s2, sum, a);
    g.mul(c2, cv, b);
    g.add(res, s2, c2);
    store_soa(g, res, Cm, ic);
}
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_mf_finish(const DevConsts *__restrict__ C_, int m, int n, int lk,
                                                   const int *__restrict__ S, const int *__restrict__ EHA,
                                                   const int *__restrict__ BHA, const int *__restrict__ ELA,
                                                   const int *__restrict__ BLA, const int *__restrict__ EHB,
                                                   const int *__restrict__ BHB, const int *__restrict__ ELB,
                                                   const int *__restrict__ BLB, SoA alpha, SoA beta, SoA Cm, int ldc) {
    alpha = rd(alpha);
    beta = rd(beta);
    Cm = rd(Cm);
    const long long e = group_id<L>();
    const long long mn = (long long)m * n;
    if (e >= mn) return;
    Grp<N, L> g(C_);
    mf_finish_rounded<N, L>(g, e, mn, (int)(e % m), (int)(e / m), lk, S, EHA, BHA, ELA, BLA, EHB, BHB, ELB, BLB,
                            alpha, beta, Cm, ldc);
}

// alpha * (band-pair sums) + beta * C of one entry, exactly in an accumulator of W limbs that
// holds every live term aligned at emin, rounded once (see k_mf_finish_exact)
template <int N, int W>
__device__ __forceinline__ void mf_exact_sum(const Grp<N, 1> &g, const DevConsts *__restrict__ C_, long long e,
                                             long long mn, const int *__restrict__ S, const MP<N> &a, const MP<N> &b,
                                             const MP<N> &cv, int live, const int (&ex)[5], int emin, MP<N> &res) {
    constexpr int WX = big_limbs<N>();
    static_assert(2 * WX + 1 <= W, "accumulator");
    // Every term is below 2^(need - 2) (k_mf_finish_exact), their sum below 2^(need - 1)
    // <= 2^(32 W - 1): one accumulator in two's complement holds every partial sum.
    Big<W> A;
    big_zero(A);
    if ((live >> 4) & 1) {  // C beta first: C and beta are not needed after it
        const Big<W> w = big_shl(
            big_widen<W>(big_mul(residues_to_big<N>(C_, cv.d), residues_to_big<N>(C_, b.d))), ex[4] - emin);
        if (cv.sign ^ b.sign)
            big_sub(A, w);
        else
            big_add(A, w);
    }
    BigN<N> Mb;
    big_zero(Mb);
    Mb.l[0] = 1u;
#pragma unroll
    for (int i = 0; i < N; i++) big_mul_add_u32(Mb, (uint32_t)C_->m[i], 0u);
    const BigN<N> Xa = residues_to_big<N>(C_, a.d);
#pragma unroll 1
    for (int t = 0; t < 4; t++) {
        if (!((live >> t) & 1)) continue;
        // the term's exponent, selected without indexing the array with t
        int ext = ex[0];
#pragma unroll
        for (int u = 1; u < 4; u++) ext = t == u ? ex[u] : ext;
        // the band pair's exact sum: residues of the signed sum, |sum| < M / 4
        int d[N];
#pragma unroll
        for (int i = 0; i < N; i++) d[i] = S[(size_t)t * N * mn + (size_t)i * mn + e];
        BigN<N> X = residues_to_big<N>(C_, d);
        int sg = big_cmp(big_shl(X, 1), Mb) >= 0;
        if (sg) {
            BigN<N> t2 = Mb;
            big_sub(t2, X);
            X = t2;
        }
        sg ^= a.sign;
        const Big<W> w = big_shl(big_widen<W>(big_mul(X, Xa)), ext - emin);
        if (sg)
            big_sub(A, w);
        else
            big_add(A, w);
    }
    const int sign = (int)(A.l[W - 1] >> 31);
    if (sign) {  // |A|
#pragma unroll
        for (int i = 0; i < W; i++) A.l[i] = ~A.l[i];
        big_add_u32(A, 1u);
    }
    if (big_is_zero(A)) {
        g.zero(res);
        return;
    }
    int exp = emin;
    big_round(A, exp, C_->mp_precision);
    const int tz = big_ctz(A);
    A = big_shr(A, tz);
    exp += tz;
    // the rounded significand has at most MP_PRECISION + 1 < log2 M bits: WX limbs hold it
    BigN<N> R;
#pragma unroll
    for (int i = 0; i < WX; i++) R.l[i] = A.l[i];
#pragma unroll
    for (int i = 0; i < N; i++) res.d[i] = big_mod_d(R, C_->md[i], C_->minv[i]);
    res.sign = sign;
    res.exp = exp;
    g.eval_full(res.d, res.ev0, res.ev1);
}

// The finish at 8 moduli with one rounding per entry: alpha * (the exact band-pair sums)
// + beta * C is formed exactly in integers (one thread per entry) and rounded once to
// MP_PRECISION bits, to nearest (ties to even); the result is that significand without
// trailing zeros, its residues and the full interval evaluation.
constexpr int MF_XS = 24, MF_XW = 48;  // 32-bit limbs of the exact accumulation (common, widest)

// One entry's terms: alpha, beta and C loaded, the live terms (bit t: band pair t, then
// C beta) with their exponents; returns the accumulator bits they need, or 0 when every
// term is zero.
template <int N>
__device__ __forceinline__ long long mf_exact_terms(const Grp<N, 1> &g, const DevConsts *__restrict__ C_, long long e,
                                                    long long mn, int row, int col, const int *__restrict__ S,
                                                    const int *__restrict__ EHA, const int *__restrict__ BLA,
                                                    const int *__restrict__ ELA, const int *__restrict__ EHB,
                                                    const int *__restrict__ BLB, const int *__restrict__ ELB,
                                                    const SoA &alpha, const SoA &beta, const SoA &Cm, long long ic,
                                                    MP<N> &a, MP<N> &b, MP<N> &cv, int &live, int (&ex)[5],
                                                    int &emin) {
    load_soa(g, a, alpha, 0);
    load_soa(g, b, beta, 0);
    load_soa(g, cv, Cm, ic);
    const bool la = BLA[row] != MF_NONE, lb = BLB[col] != MF_NONE;
    auto nonzero = [](const int *d) {
        int o = 0;
#pragma unroll
        for (int i = 0; i < N; i++) o |= d[i];
        return o != 0;
    };
    const bool anz = nonzero(a.d);
    live = 0;
#pragma unroll
    for (int p = 0; p < 4; p++) {
        const int ba_ = p >> 1, bb_ = p & 1;
        if (anz && !((!ba_ && !la) || (!bb_ && !lb))) {
            int d[N];
#pragma unroll
            for (int i = 0; i < N; i++) d[i] = S[(size_t)p * N * mn + (size_t)i * mn + e];
            live |= nonzero(d) ? 1 << p : 0;
        }
        ex[p] = (ba_ ? EHA[row] : ELA[row]) + (bb_ ? EHB[col] : ELB[col]) + a.exp;
    }
    live |= nonzero(b.d) && nonzero(cv.d) ? 1 << 4 : 0;
    ex[4] = cv.exp + b.exp;
    emin = 0x7fffffff;
    int emax = -0x7fffffff;
#pragma unroll
    for (int t = 0; t < 5; t++)
        if ((live >> t) & 1) {
            emin = ex[t] < emin ? ex[t] : emin;
            emax = ex[t] > emax ? ex[t] : emax;
        }
    if (!live) return 0;
    // a product of two numbers below M has at most 2 log2 M + 2 bits
    return (long long)emax - emin + 2 * (C_->log2m + 1) + 2;
}

// The common case: entries whose aligned terms fit MF_XS limbs. The others are appended
// to a list (wcnt, wlist) for k_mf_finish_wide, and their C entries are left untouched.
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_mf_finish_exact(const DevConsts *__restrict__ C_, int m, int n,
                                                         const int *__restrict__ S, const int *__restrict__ EHA,
                                                         const int *__restrict__ ELA, const int *__restrict__ BLA,
                                                         const int *__restrict__ EHB, const int *__restrict__ ELB,
                                                         const int *__restrict__ BLB, SoA alpha, SoA beta, SoA Cm,
                                                         int ldc, int *wcnt, int *wlist) {
    alpha = rd(alpha);
    beta = rd(beta);
    Cm = rd(Cm);
    const long long e = (long long)blockIdx.x * BLK + threadIdx.x;
    const long long mn = (long long)m * n;
    if (e >= mn) return;
    const int row = (int)(e % m), col = (int)(e / m);
    const long long ic = (long long)col * ldc + row;
    Grp<N, 1> g(C_);
    MP<N> a, b, cv, res;
    int live, ex[5], emin;
    const long long need =
        mf_exact_terms<N>(g, C_, e, mn, row, col, S, EHA, BLA, ELA, EHB, BLB, ELB, alpha, beta, Cm, ic, a, b, cv, live,
                          ex, emin);
    if (need > 32 * MF_XS) {
        wlist[atomicAdd(wcnt, 1)] = (int)e;
        return;
    }
    if (need == 0)
        g.zero(res);
    else
        mf_exact_sum<N, MF_XS>(g, C_, e, mn, S, a, b, cv, live, ex, emin, res);
    store_soa(g, res, Cm, ic);
}

// The listed entries: exactly in MF_XW limbs, or, when the terms span more than that, the
// rounded finish (mf_finish_rounded). Grid-stride over the list.
template <int N, int L>
__global__ __launch_bounds__(BLK) void k_mf_finish_wide(const DevConsts *__restrict__ C_, int m, int n, int lk,
                                                        const int *__restrict__ S, const int *__restrict__ EHA,
                                                        const int *__restrict__ BHA, const int *__restrict__ ELA,
                                                        const int *__restrict__ BLA, const int *__restrict__ EHB,
                                                        const int *__restrict__ BHB, const int *__restrict__ ELB,
                                                        const int *__restrict__ BLB, SoA alpha, SoA beta, SoA Cm,
                                                        int ldc, const int *__restrict__ wcnt,
                                                        const int *__restrict__ wlist) {
    alpha = rd(alpha);
    beta = rd(beta);
    Cm = rd(Cm);
    const long long mn = (long long)m * n;
    const int cnt = *wcnt;
    Grp<N, 1> g(C_);
    for (int w = blockIdx.x * BLK + threadIdx.x; w < cnt; w += gridDim.x * BLK) {
        const long long e = wlist[w];
        const int row = (int)(e % m), col = (int)(e / m);
        const long long ic = (long long)col * ldc + row;
        MP<N> a, b, cv, res;
        int live, ex[5], emin;
        const long long need =
            mf_exact_terms<N>(g, C_, e, mn, row, col, S, EHA, BLA, ELA, EHB, BLB, ELB, alpha, beta, Cm, ic, a, b, cv,
                              live, ex, emin);
        if (need <= 32 * MF_XW) {
            mf_exact_sum<N, MF_XW>(g, C_, e, mn, S, a, b, cv, live, ex, emin, res);
            store_soa(g, res, Cm, ic);
        } else {
            mf_finish_rounded<N, 1>(g, e, mn, row, col, lk, S, EHA, BHA, ELA, BLA, EHB, BHB, ELB, BLB, alpha, beta, Cm,
                                    ldc);
        }
    }
}

}  // namespace mph
