// Host-side launchers of the kernel templates for one (N, L) pair, exposed
// as a table of plain function pointers so that api.cpp (the C ABI) can
// dispatch at run time on the context's moduli count and the chosen lanes
// per number.
#pragma once

#include <hip/hip_runtime.h>

#include "ops.hpp"

#if defined(__HIPCC__) && defined(MPH_KERNELS)
#include "kernels.hpp"
#include "kernels_mfma.hpp"
#include "kernels_batch.hpp"
#include "kernels_blas.hpp"

namespace mph {

inline unsigned grid_for(long long groups, int L) {
    const long long threads = groups * L;
    return (unsigned)((threads + BLK - 1) / BLK);
}

template <int N, int L>
struct Launch {
    static hipError_t set_d(const DevConsts *c, const double *x, int count, SoA so, char *aos, bool soa,
                            hipStream_t st) {
        if (count <= 0) return hipSuccess;
        if (soa)
            hipLaunchKernelGGL((k_set_d<N, L, true>), dim3(grid_for(count, L)), dim3(BLK), 0, st, c, x, count, so,
                               AoSRef<N>{aos});
        else
            hipLaunchKernelGGL((k_set_d<N, L, false>), dim3(grid_for(count, L)), dim3(BLK), 0, st, c, x, count, so,
                               AoSRef<N>{aos});
        return hipGetLastError();
    }
    static hipError_t elem(const DevConsts *c, int op, int count, const char *x, const char *y, const double *yd,
                           char *r, hipStream_t st) {
        if (count <= 0) return hipSuccess;
        const dim3 gr(grid_for(count, L)), bl(BLK);
        AoSRef<N> X{(char *)x}, Y{(char *)y}, Rr{r};
        switch (op) {
            case 0: hipLaunchKernelGGL((k_elem<N, L, 0>), gr, bl, 0, st, c, count, X, Y, yd, Rr); break;
            case 1: hipLaunchKernelGGL((k_elem<N, L, 1>), gr, bl, 0, st, c, count, X, Y, yd, Rr); break;
            case 2: hipLaunchKernelGGL((k_elem<N, L, 2>), gr, bl, 0, st, c, count, X, Y, yd, Rr); break;
            case 3: hipLaunchKernelGGL((k_elem<N, L, 3>), gr, bl, 0, st, c, count, X, Y, yd, Rr); break;
            case 4: hipLaunchKernelGGL((k_elem<N, L, 4>), gr, bl, 0, st, c, count, X, Y, yd, Rr); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    static hipError_t exp_out(const DevConsts *c, int mode, int count, const char *x, SoA so, bool soa, double *out,
                              unsigned *limbs, int *sg, int *ex, hipStream_t st) {
        if (count <= 0) return hipSuccess;
        const dim3 gr(grid_for(count, 1)), bl(BLK);
        uint32_t *lb = (uint32_t *)limbs;
        if (soa) {
            if (mode == 0)
                hipLaunchKernelGGL((k_export<N, L, SoA, 0>), gr, bl, 0, st, c, count, so, out, lb, sg, ex);
            else
                hipLaunchKernelGGL((k_export<N, L, SoA, 1>), gr, bl, 0, st, c, count, so, out, lb, sg, ex);
        } else {
            const AoSRef<N> X{(char *)x};
            if (mode == 0)
                hipLaunchKernelGGL((k_export<N, L, AoSRef<N>, 0>), gr, bl, 0, st, c, count, X, out, lb, sg, ex);
            else
                hipLaunchKernelGGL((k_export<N, L, AoSRef<N>, 1>), gr, bl, 0, st, c, count, X, out, lb, sg, ex);
        }
        return hipGetLastError();
    }
    static hipError_t cmp(const DevConsts *c, int mode, int count, const char *x, const char *y, const int *dx,
                          const int *dy, int *out, hipStream_t st) {
        if (count <= 0) return hipSuccess;
        const dim3 gr(grid_for(count, L)), bl(BLK);
        AoSRef<N> X{(char *)x}, Y{(char *)y};
        switch (mode) {
            case 0: hipLaunchKernelGGL((k_cmp<N, L, 0>), gr, bl, 0, st, c, count, X, Y, dx, dy, out); break;
            case 1: hipLaunchKernelGGL((k_cmp<N, L, 1>), gr, bl, 0, st, c, count, X, Y, dx, dy, out); break;
            case 2: hipLaunchKernelGGL((k_cmp<N, L, 2>), gr, bl, 0, st, c, count, X, Y, dx, dy, out); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    // pass 0: the main kernel; pass 1: the fixup kernel over the chains the main
    // kernel marked (16 and 32 moduli; nothing to do over 8 moduli)
    static constexpr bool SPLIT = !Grp<N, L>::INL;
    template <int MODE>
    static void fold_pass(int pass, const dim3 &gr, const DevConsts *c, long long cnt, long long nvt, SoA xs, int incx,
                          SoA ys, int incy, AoSRef<N> XA, AoSRef<N> YA, AoSRef<N> O, hipStream_t st) {
        const dim3 bl(BLK);
        if (pass == 0)
            hipLaunchKernelGGL((k_fold<N, L, MODE, !SPLIT, false>), gr, bl, 0, st, c, cnt, nvt, xs, incx, ys, incy, XA,
                               YA, O);
        else if constexpr (SPLIT)
            hipLaunchKernelGGL((k_fold<N, L, MODE, true, true>), gr, bl, 0, st, c, cnt, nvt, xs, incx, ys, incy, XA,
                               YA, O);
    }
    static hipError_t fold(const DevConsts *c, int mode, int pass, long long cnt, long long nvt, SoA xs, int incx,
                           SoA ys, int incy, const char *xa, const char *ya, char *out, hipStream_t st) {
        if (nvt <= 0) return hipSuccess;
        const dim3 gr(grid_for(nvt, L)), bl(BLK);
        AoSRef<N> XA{(char *)xa}, YA{(char *)ya}, O{out};
        if (mode == 0)
            fold_pass<0>(pass, gr, c, cnt, nvt, xs, incx, ys, incy, XA, YA, O, st);
        else if (mode == 1)
            fold_pass<1>(pass, gr, c, cnt, nvt, xs, incx, ys, incy, XA, YA, O, st);
        else if (pass == 0)
            hipLaunchKernelGGL((k_fold<N, L, 2, true, false>), gr, bl, 0, st, c, cnt, nvt, xs, incx, ys, incy, XA, YA,
                               O);
        return hipGetLastError();
    }
    static hipError_t tree(const DevConsts *c, char *part, int b, unsigned p2, int nblocks, char *out, hipStream_t st) {
        if (nblocks <= 0) return hipSuccess;
        hipLaunchKernelGGL((k_tree<N, L>), dim3(nblocks), dim3(BLK), 0, st, c, AoSRef<N>{part}, b, p2, AoSRef<N>{out});
        return hipGetLastError();
    }
    static hipError_t aos_to_soa(const DevConsts *c, const char *src, SoA dst, int count, hipStream_t st) {
        if (count <= 0) return hipSuccess;
        hipLaunchKernelGGL((k_aos_to_soa<N, L>), dim3(grid_for(count, L)), dim3(BLK), 0, st, c,
                           AoSRef<N>{(char *)src}, dst, count);
        return hipGetLastError();
    }
    static hipError_t soa_to_aos(const DevConsts *c, SoA src, char *dst, int count, hipStream_t st) {
        if (count <= 0) return hipSuccess;
        hipLaunchKernelGGL((k_soa_to_aos<N, L>), dim3(grid_for(count, L)), dim3(BLK), 0, st, c, src, AoSRef<N>{dst},
                           count);
        return hipGetLastError();
    }
    static hipError_t axpy(const DevConsts *c, int n, SoA alpha, SoA x, int incx, SoA y, int incy, hipStream_t st) {
        if (n <= 0) return hipSuccess;
        hipLaunchKernelGGL((k_axpy<N, L>), dim3(grid_for(n, L)), dim3(BLK), 0, st, c, n, alpha, x, incx, y, incy);
        return hipGetLastError();
    }
    static hipError_t axpy_v2(const DevConsts *c, int n, const char *alpha, const char *x, const char *y, char *r,
                              bool neg, hipStream_t st) {
        if (n <= 0) return hipSuccess;
        const dim3 gr(grid_for(n, L)), bl(BLK);
        const AoSRef<N> A{(char *)alpha}, X{(char *)x}, Y{(char *)y}, R{r};
        if (neg)
            hipLaunchKernelGGL((k_axpy_v2<N, L, true>), gr, bl, 0, st, c, n, A, X, Y, R);
        else
            hipLaunchKernelGGL((k_axpy_v2<N, L, false>), gr, bl, 0, st, c, n, A, X, Y, R);
        return hipGetLastError();
    }
    static hipError_t hostop(const DevConsts *c, int op, int count, const char *x, const char *y, char *r, double *out,
                             hipStream_t st) {
        if (count <= 0) return hipSuccess;
        const dim3 gr(grid_for(count, 1)), bl(BLK);
        const AoSRef<N> X{(char *)x}, Y{(char *)y}, R{r};
        if (op == 0)
            hipLaunchKernelGGL((k_div<N, L>), gr, bl, 0, st, c, count, X, Y, R);
        else if (op == 1)
            hipLaunchKernelGGL((k_sqrt<N, L, 0>), gr, bl, 0, st, c, count, X, R, out);
        else
            hipLaunchKernelGGL((k_sqrt<N, L, 1>), gr, bl, 0, st, c, count, X, R, out);
        return hipGetLastError();
    }
    static hipError_t scal(const DevConsts *c, int len, SoA s, SoA v, int inc, char *out, bool inplace,
                           hipStream_t st) {
        if (len <= 0) return hipSuccess;
        if (inplace)
            hipLaunchKernelGGL((k_scal<N, L, true>), dim3(grid_for(len, L)), dim3(BLK), 0, st, c, len, s, v, inc,
                               AoSRef<N>{out});
        else
            hipLaunchKernelGGL((k_scal<N, L, false>), dim3(grid_for(len, L)), dim3(BLK), 0, st, c, len, s, v, inc,
                               AoSRef<N>{out});
        return hipGetLastError();
    }
    // gemv passes: 0 preparation (split form: row minima of A's exponents, minimum
    // and pre-scaled residues of bx), 1 main kernel (split form: the residues), 2 split
    // form: the scalar chains (independent of pass 1, so a caller may run the two on two
    // streams), 3 split form: the residues moved to the chains' results, then the fixup of
    // marked partials. aux: the split form's bxp (N*n ints), EA (m ints), EB (1 int).
    static bool gemv_split(bool trans, int b3) { return SPLIT && !trans && b3 <= 65535; }
    static size_t gemv_aux(bool trans, int m, int n) {
        return (SPLIT && !trans) ? 4 * ((size_t)N * n + m + 1) : 0;
    }
    template <bool TR>
    static void gemv_pass(int pass, const DevConsts *c, int m, int n, SoA A, int lda, AoSRef<N> bx, int b3,
                          long long nvt, AoSRef<N> part, char *aux, hipStream_t st) {
        const dim3 gr(grid_for(nvt, L)), bl(BLK);
        const bool split = gemv_split(TR, b3);
        int *bxp = (int *)aux, *EA = bxp + (size_t)N * n, *EB = EA + m;
        if (pass == 0) {
            if (split) {
                const unsigned kc = (unsigned)((n + EMIN_K - 1) / EMIN_K);
                hipLaunchKernelGGL((k_fill_int<N, L>), dim3(grid_for((long long)m + 1, 1)), bl, 0, st, EA, m + 1,
                                   0x7fffffff);
                hipLaunchKernelGGL((k_gemm_emin<N, L, SoA, 0>), dim3((m + EMIN_J - 1) / EMIN_J, kc), bl, 0, st, m, n, 0,
                                   A, lda, EA);
                hipLaunchKernelGGL((k_gemm_emin<N, L, AoSRef<N>, 1>), dim3(1, kc), bl, 0, st, 1, n, 0, bx, 0, EB);
                hipLaunchKernelGGL((k_gemm_prescale<N, L, AoSRef<N>, 1>), dim3(grid_for(n, 1)), bl, 0, st, c, 1, n, 0,
                                   bx, 0, EB, bxp);
            }
        } else if (pass == 1) {
            if (split) {
                constexpr int RW = GV_RT * GR / N;  // rows per workgroup of k_gemv_resid
                hipLaunchKernelGGL((k_gemv_resid<N, L>), dim3((m + RW - 1) / RW, b3), dim3(GV_RT), 0, st, c, m, n, A,
                                   lda, bxp, EA, EB, b3, part);
            }
            else if (!TR && b3 <= 65535)  // grid y limit
                hipLaunchKernelGGL((k_gemv_rows<N, L, !SPLIT>), dim3((m + BLK / L - 1) / (BLK / L), b3), bl, 0, st, c,
                                   m, n, A, lda, bx, b3, part);
            else
                hipLaunchKernelGGL((k_gemv_fold<N, L, TR, !SPLIT, false>), gr, bl, 0, st, c, m, n, A, lda, bx, b3,
                                   nvt, part);
        } else if (pass == 2) {
            if (split) {
                const dim3 gc(grid_for((long long)m * b3, 1));
                const bool narrow = (long long)lda * n <= (1LL << 28);
                if (narrow && m % 64 == 0)
                    hipLaunchKernelGGL((k_gemv_chain<N, L, true, true>), gc, bl, 0, st, c, m, n, A, lda, bx, b3, part);
                else if (narrow)
                    hipLaunchKernelGGL((k_gemv_chain<N, L, true, false>), gc, bl, 0, st, c, m, n, A, lda, bx, b3, part);
                else
                    hipLaunchKernelGGL((k_gemv_chain<N, L, false, false>), gc, bl, 0, st, c, m, n, A, lda, bx, b3,
                                       part);
            }
        } else if constexpr (SPLIT) {
            if (split)
                hipLaunchKernelGGL((k_gemv_scale<N, L>), dim3(grid_for(nvt * (N / GR), 1)), bl, 0, st, c, nvt, b3,
                                   (const int *)EA, (const int *)EB, part);
            hipLaunchKernelGGL((k_gemv_fold<N, L, TR, true, true>), gr, bl, 0, st, c, m, n, A, lda, bx, b3, nvt, part);
        }
    }
    static hipError_t gemv_fold(const DevConsts *c, bool trans, int pass, int m, int n, SoA A, int lda, const char *bx,
                                int b3, long long nvt, char *part, char *aux, hipStream_t st) {
        if (nvt <= 0) return hipSuccess;
        if (trans)
            gemv_pass<true>(pass, c, m, n, A, lda, AoSRef<N>{(char *)bx}, b3, nvt, AoSRef<N>{part}, aux, st);
        else
            gemv_pass<false>(pass, c, m, n, A, lda, AoSRef<N>{(char *)bx}, b3, nvt, AoSRef<N>{part}, aux, st);
        return hipGetLastError();
    }
    static hipError_t gemv_final(const DevConsts *c, int len, SoA y, int incy, const char *sums, hipStream_t st) {
        if (len <= 0) return hipSuccess;
        hipLaunchKernelGGL((k_gemv_final<N, L>), dim3(grid_for(len, L)), dim3(BLK), 0, st, c, len, y, incy,
                           AoSRef<N>{(char *)sums});
        return hipGetLastError();
    }
    // gemm passes. Over 16 and 32 moduli (split form, see kernels.hpp):
    //   0 exponent minima and the int8 limb planes of the pre-scaled residues, 1 scalar chains,
    //   2 residue products (INT8 MFMA) and digits, 3 fixup of the marked chains, 4 epilogue.
    // Over 8 moduli: 1 full sums, 4 epilogue. bits: bit length of the largest modulus.
    // ws: m*n sums, then (split form) the residue sums S (N*m*n ints), the minima (m + n
    // ints) and the limb planes of op(A) and op(B).
    static constexpr int GEMM_NL = 5;  // limbs per residue reserved (moduli below 2^35)
    static int gemm_limbs(int bits) { return bits <= 4 * LIMB_BITS ? 4 : 5; }
    static size_t gemm_ws(int m, int n, int k) {
        const size_t sums = ((size_t)m * n * (4 * N + 40) + 255) / 256 * 256;
        if (!SPLIT) return sums;
        const size_t ints = ((4 * ((size_t)N * m * n + m + n)) + 255) / 256 * 256;
        const size_t MPd = (m + 63) / 64 * 64, NPd = (n + 63) / 64 * 64,
                     KPd = (size_t)((k > 0 ? k : 1) + LIMB_T - 1) / LIMB_T * LIMB_T;
        return sums + ints + (size_t)N * GEMM_NL * (MPd + NPd) * KPd;
    }
    // Passes of the split-form gemm: 0 = exponent minima, int8 limb planes and the INT8 MFMA
    // residue products; 1 = the scalar chains; 2 = digits from both (needs 0 and 1); 3 = the
    // fixup of REDO chains; 4 = the alpha/beta epilogue. Passes 0 and 1 are independent, so a
    // caller may run pass 0 on another stream. Over 8 moduli only pass 1 (the lane-group
    // reduction) and pass 4 do work.
    template <class VM>
    static void gemm_pass(int pass, const DevConsts *c, int bits, int ta, int tb, int m, int n, int k, VM A, int lda,
                          VM B, int ldb, char *ws, hipStream_t st) {
        const long long mn = (long long)m * n;
        const dim3 gr(grid_for(mn, L)), bl(BLK);
        const AoSRef<N> sums{ws};
        const size_t sb_ = ((size_t)mn * (4 * N + 40) + 255) / 256 * 256;
        int *S = (int *)(ws + sb_), *EA = S + (size_t)N * mn, *EB = EA + m;
        const size_t ib_ = ((4 * ((size_t)N * mn + m + n)) + 255) / 256 * 256;
        const int MP = (m + 63) / 64 * 64, NP = (n + 63) / 64 * 64,
                  KP = ((k > 0 ? k : 1) + LIMB_T - 1) / LIMB_T * LIMB_T;
        const int nl = gemm_limbs(bits);
        signed char *PA = (signed char *)(ws + sb_ + ib_), *PB = PA + (size_t)N * nl * MP * KP;
        if constexpr (SPLIT) {
            if (pass == 0) {
                hipLaunchKernelGGL((k_fill_int<N, L>), dim3(grid_for((long long)m + n, 1)), bl, 0, st, EA, m + n,
                                   k > 0 ? 0x7fffffff : 0);
                if (k > 0) {
                    const unsigned kc = (unsigned)((k + EMIN_K - 1) / EMIN_K);
                    hipLaunchKernelGGL((k_gemm_emin<N, L, VM, 0>), dim3((m + EMIN_J - 1) / EMIN_J, kc), bl, 0, st, m,
                                       k, ta, A, lda, EA);
                    hipLaunchKernelGGL((k_gemm_emin<N, L, VM, 1>), dim3((n + EMIN_J - 1) / EMIN_J, kc), bl, 0, st, n,
                                       k, tb, B, ldb, EB);
                }
                hipLaunchKernelGGL((k_gemm_limbs<N, L, VM, 0>), dim3(grid_for((long long)MP * KP, 1)), bl, 0, st, c, m,
                                   MP, k, KP, nl, ta, A, lda, EA, PA);
                hipLaunchKernelGGL((k_gemm_limbs<N, L, VM, 1>), dim3(grid_for((long long)NP * KP, 1)), bl, 0, st, c, n,
                                   NP, k, KP, nl, tb, B, ldb, EB, PB);
                mf_products(c, nl, m, n, MP, NP, KP, PA, PB, nullptr, nullptr, S, st);
            } else if (pass == 1) {
                // no transposes and element offsets below 2^28: the specialised chain kernel
                if constexpr (std::is_same<VM, SoA>::value) {
                    if (ta == 0 && tb == 0 && (long long)lda * k <= (1LL << 28) && (long long)ldb * n <= (1LL << 28)) {
                        if (m % 64 == 0)
                            hipLaunchKernelGGL((k_gemm_chain_nn<N, L, true>), dim3(grid_for(mn, 1)), bl, 0, st, c, m,
                                               n, k, A, lda, B, ldb, sums);
                        else
                            hipLaunchKernelGGL((k_gemm_chain_nn<N, L, false>), dim3(grid_for(mn, 1)), bl, 0, st, c, m,
                                               n, k, A, lda, B, ldb, sums);
                        return;
                    }
                }
                hipLaunchKernelGGL((k_gemm_chain<N, L, VM>), dim3(grid_for(mn, 1)), bl, 0, st, c, ta, tb, m, n, k, A,
                                   lda, B, ldb, sums);
            } else if (pass == 2) {
                hipLaunchKernelGGL((k_gemm_digits<N, L>), dim3(grid_for(mn, 1)), bl, 0, st, c, m, n, S, EA, EB, sums);
            } else if (pass == 3) {
                hipLaunchKernelGGL((k_gemm_sum<N, L, true, true, VM>), gr, bl, 0, st, c, ta, tb, m, n, k, A, lda, B,
                                   ldb, sums);
            }
        } else if (pass == 1) {
            hipLaunchKernelGGL((k_gemm_sum<N, L, true, false, VM>), gr, bl, 0, st, c, ta, tb, m, n, k, A, lda, B, ldb,
                               sums);
        }
    }
    static hipError_t gemm(const DevConsts *c, int pass, int bits, int m, int n, int k, SoA alpha, SoA A, int lda,
                           SoA B, int ldb, SoA beta, SoA Cm, int ldc, char *ws, hipStream_t st) {
        if ((long long)m * n <= 0) return hipSuccess;
        if (pass < 4)
            gemm_pass<SoA>(pass, c, bits, 0, 0, m, n, k, A, lda, B, ldb, ws, st);
        else
            hipLaunchKernelGGL((k_gemm_fin<N, L, false, SoA>), dim3(grid_for((long long)m * n, L)), dim3(BLK), 0, st,
                               c, m, n, alpha, beta, Cm, ldc, AoSRef<N>{ws});
        return hipGetLastError();
    }
    static hipError_t gemm_v2(const DevConsts *c, int pass, int bits, int ta, int tb, int m, int n, int k,
                              const char *alpha, const char *A, int lda, const char *B, int ldb, const char *beta,
                              char *Cm, int ldc, char *ws, hipStream_t st) {
        if ((long long)m * n <= 0) return hipSuccess;
        if (pass < 4)
            gemm_pass<AoSRef<N>>(pass, c, bits, ta, tb, m, n, k, AoSRef<N>{(char *)A}, lda, AoSRef<N>{(char *)B}, ldb,
                                 ws, st);
        else
            hipLaunchKernelGGL((k_gemm_fin<N, L, true, AoSRef<N>>), dim3(grid_for((long long)m * n, L)), dim3(BLK),
                               0, st, c, m, n, AoSRef<N>{(char *)alpha}, AoSRef<N>{(char *)beta}, AoSRef<N>{Cm}, ldc,
                               AoSRef<N>{ws});
        return hipGetLastError();
    }
    // ---- precision form on INT8 MFMA (kernels_mfma.hpp) ----
    static constexpr long long rup(long long v, long long q) { return (v + q - 1) / q * q; }
    // ws: S (4 band pairs x N*m*n ints); per row and column: minimum exponent, top bit, top
    // bit of the low band, high / low band exponents and bit bounds (7 (m + n) ints); then the
    // limb planes of both bands of op(A) and op(B)
    static size_t mf_ints(int m, int n) { return 4 * (size_t)N * m * n + 7 * ((size_t)m + n); }
    static size_t gemm_mfma_ws(int bits, int m, int n, int k) {
        const int nl = gemm_limbs(bits);
        const size_t MPd = rup(m, 64), NPd = rup(n, 64), KPd = rup(k > 0 ? k : 1, MF_T);
        return (4 * mf_ints(m, n) + 255) / 256 * 256 + 2 * (size_t)N * nl * (MPd + NPd) * KPd + mf_wide_bytes(m, n);
    }
    // 8 moduli: the count and list of the entries of the wide exact finish (k_mf_finish_wide)
    static size_t mf_wide_bytes(int m, int n) { return N == 8 ? 4 * (64 + (size_t)m * n) : 0; }
    // residue products of limb planes on INT8 MFMA (k_mf_gemm), one launch
    static void mf_products(const DevConsts *c, int nl, int m, int n, int MP, int NP, int KP, const signed char *pa,
                            const signed char *pb, const int *fa, const int *fb, int *S, hipStream_t st) {
        const dim3 gg(NP / 64, MP / 64, N), bl(BLK);
        const bool ch = KP > MF_KONE;
        if (nl == 4 && !ch)
            hipLaunchKernelGGL((k_mf_gemm<N, L, 4, false>), gg, bl, 0, st, c, m, n, MP, NP, KP, pa, pb, fa, fb, S);
        else if (nl == 4)
            hipLaunchKernelGGL((k_mf_gemm<N, L, 4, true>), gg, bl, 0, st, c, m, n, MP, NP, KP, pa, pb, fa, fb, S);
        else if (!ch)
            hipLaunchKernelGGL((k_mf_gemm<N, L, 5, false>), gg, bl, 0, st, c, m, n, MP, NP, KP, pa, pb, fa, fb, S);
        else
            hipLaunchKernelGGL((k_mf_gemm<N, L, 5, true>), gg, bl, 0, st, c, m, n, MP, NP, KP, pa, pb, fa, fb, S);
    }
    static hipError_t gemm_mfma(const DevConsts *c, int bits, int log2m, int ta, int tb, int m, int n, int k,
                                SoA alpha, SoA A, int lda, SoA B, int ldb, SoA beta, SoA Cm, int ldc, char *ws,
                                hipStream_t st) {
        if ((long long)m * n <= 0) return hipSuccess;
        const int nl = gemm_limbs(bits);
        const int MP = (int)rup(m, 64), NP = (int)rup(n, 64), KP = (int)rup(k > 0 ? k : 1, MF_T);
        int *S = (int *)ws;
        int *V = S + 4 * (size_t)N * m * n;  // 7 arrays of m + n: rows of op(A) first, then columns of op(B)
        const size_t mn = (size_t)m + n;
        int *EMIN = V, *TOP = V + mn, *TOPL = V + 2 * mn, *EH = V + 3 * mn, *BH = V + 4 * mn, *EL = V + 5 * mn,
            *BL = V + 6 * mn;
        signed char *PA = (signed char *)ws + (4 * mf_ints(m, n) + 255) / 256 * 256;
        const size_t bandA = (size_t)N * nl * MP * KP, bandB = (size_t)N * nl * NP * KP;
        signed char *PB = PA + 2 * bandA;
        const dim3 bl(BLK);
        hipLaunchKernelGGL((k_fill_int<N, L>), dim3(grid_for((long long)mn, 1)), bl, 0, st, EMIN, (int)mn, 0x7fffffff);
        hipLaunchKernelGGL((k_fill_int<N, L>), dim3(grid_for(2 * (long long)mn, 1)), bl, 0, st, TOP, (int)(2 * mn),
                           MF_NONE);
        const unsigned kc = (unsigned)((k + EMIN_K - 1) / EMIN_K);
        const dim3 ga((m + EMIN_J - 1) / EMIN_J, kc), gb((n + EMIN_J - 1) / EMIN_J, kc);
        if (k > 0) {
            hipLaunchKernelGGL((k_gemm_emin<N, L, SoA, 0, true>), ga, bl, 0, st, m, k, ta, A, lda, EMIN);
            hipLaunchKernelGGL((k_gemm_emin<N, L, SoA, 1, true>), gb, bl, 0, st, n, k, tb, B, ldb, EMIN + m);
            hipLaunchKernelGGL((k_mf_top<N, L, SoA, 0, false>), ga, bl, 0, st, c, m, k, ta, A, lda, nullptr, TOP);
            hipLaunchKernelGGL((k_mf_top<N, L, SoA, 1, false>), gb, bl, 0, st, c, n, k, tb, B, ldb, nullptr, TOP + m);
        }
        int lk = 0;
        while ((1LL << lk) < (long long)(k > 0 ? k : 1)) lk++;
        const int half = (log2m - 2 - lk) / 2;
        const dim3 gv(grid_for((long long)mn, 1));
        hipLaunchKernelGGL((k_mf_band<N, L>), gv, bl, 0, st, (int)mn, half, EMIN, TOP, EH, BH);
        if (k > 0) {
            hipLaunchKernelGGL((k_mf_top<N, L, SoA, 0, true>), ga, bl, 0, st, c, m, k, ta, A, lda, EH, TOPL);
            hipLaunchKernelGGL((k_mf_top<N, L, SoA, 1, true>), gb, bl, 0, st, c, n, k, tb, B, ldb, EH + m, TOPL + m);
        }
        hipLaunchKernelGGL((k_mf_band<N, L>), gv, bl, 0, st, (int)mn, half, EMIN, TOPL, EL, BL);
        hipLaunchKernelGGL((k_mf_prescale<N, L, SoA, 0>), dim3(grid_for((long long)MP * KP, 1)), bl, 0, st, c, m, MP,
                           k, KP, nl, ta, A, lda, EH, EL, BL, PA);
        hipLaunchKernelGGL((k_mf_prescale<N, L, SoA, 1>), dim3(grid_for((long long)NP * KP, 1)), bl, 0, st, c, n, NP,
                           k, KP, nl, tb, B, ldb, EH + m, EL + m, BL + m, PB);
        for (int p = 0; p < 4; p++) {
            const int ba = p >> 1, bb = p & 1;
            const signed char *pa = PA + (size_t)ba * bandA, *pb = PB + (size_t)bb * bandB;
            const int *fa = ba ? nullptr : BL, *fb = bb ? nullptr : BL + m;
            mf_products(c, nl, m, n, MP, NP, KP, pa, pb, fa, fb, S + (size_t)p * N * m * n, st);
        }
        if constexpr (N == 8) {  // one rounding per entry (exact integer epilogue)
            int *wcnt = (int *)(PB + 2 * bandB), *wlist = wcnt + 64;
            hipLaunchKernelGGL((k_fill_int<N, L>), dim3(1), bl, 0, st, wcnt, 1, 0);
            hipLaunchKernelGGL((k_mf_finish_exact<N, L>), dim3(grid_for((long long)m * n, 1)), bl, 0, st, c, m, n, S,
                               EH, EL, BL, EH + m, EL + m, BL + m, alpha, beta, Cm, ldc, wcnt, wlist);
            const long long gw = grid_for((long long)m * n, 1);
            hipLaunchKernelGGL((k_mf_finish_wide<N, L>), dim3((unsigned)(gw < 1024 ? gw : 1024)), bl, 0, st, c, m, n,
                               lk, S, EH, BH, EL, BL, EH + m, BH + m, EL + m, BL + m, alpha, beta, Cm, ldc,
                               (const int *)wcnt, (const int *)wlist);
        } else
            hipLaunchKernelGGL((k_mf_finish<N, L>), dim3(grid_for((long long)m * n, L)), bl, 0, st, c, m, n, lk, S, EH,
                               BH, EL, BL, EH + m, BH + m, EL + m, BL + m, alpha, beta, Cm, ldc);
        return hipGetLastError();
    }
    // workgroups of k_spmv_resid_wave: BLK / SPW_T rows each per sweep, at most 16384
    static unsigned rows_blocks(int m) {
        const long long b = ((long long)m + BLK / SPW_T - 1) / (BLK / SPW_T);
        return (unsigned)(b < 16384 ? b : 16384);
    }
    // Passes of the split form (16 and 32 moduli): 0 the rows' order (CSR rows grouped by
    // length, longest first) and the scalar chains (and, but for CSR, the residues); 2 the
    // residues of CSR's short rows, 3 those of its long rows (2 and 3 take disjoint rows, so
    // they may run at once on two streams); 1 the rows the chains marked REDO, recomputed by
    // the fused kernel. Without the split form (8 moduli) pass 0 is the fused kernel and the
    // others do nothing. ws (split form, spmv_ws_bytes(m)): the marked-row count, class
    // counts, offsets and cursors, then the row order and the marked-row list.
    template <int FMT>
    static void spmv_fmt(const DevConsts *c, int pass, int m, const SpMat &A, const char *x, char *y, int tpr,
                         int *ws, hipStream_t st) {
        const dim3 gr(grid_for((long long)m * tpr, L)), bl(BLK);
        const AoSRef<N> X{(char *)x}, Y{y};
        int *rcnt = ws, *hist = ws + 1, *off = hist + SP_NB, *cur = off + SP_NB;
        int *perm = cur + SP_NB, *rlist = perm + m;
        const int *nlong = off + SP_LONG_CLASS;  // the long rows lead perm
        if constexpr (SPLIT) {
            if (pass == 0) {
                hipLaunchKernelGGL((k_fill_int<N, L>), dim3(1), bl, 0, st, ws, 1 + 3 * SP_NB, 0);
                const int *pm = nullptr;
                if (FMT == SP_CSR) {
                    const dim3 gm(
                        (unsigned)(((long long)m + (long long)BLK * SP_RPT - 1) / ((long long)BLK * SP_RPT)));
                    hipLaunchKernelGGL((k_sp_bin_count<N, L>), gm, bl, 0, st, m, A.ptr, hist);
                    hipLaunchKernelGGL((k_sp_bin_scan<N, L>), dim3(1), dim3(64), 0, st, (const int *)hist, off);
                    hipLaunchKernelGGL((k_sp_bin_scatter<N, L>), gm, bl, 0, st, m, A.ptr, (const int *)off, cur,
                                       perm);
                    pm = perm;
                }
                if (FMT == SP_CSR && tpr == 1)  // long serial chains: a deeper ring of gathers
                    hipLaunchKernelGGL((k_spmv_chain<N, FMT, SP_PF1>), dim3(grid_for((long long)m, 1)), bl, 0, st,
                                       c, m, A, X, Y, tpr, pm, rcnt, rlist);
                else
                    hipLaunchKernelGGL((k_spmv_chain<N, FMT>), dim3(grid_for((long long)m * tpr, 1)), bl, 0, st, c,
                                       m, A, X, Y, tpr, pm, rcnt, rlist);
                if (FMT != SP_CSR)
                    hipLaunchKernelGGL((k_spmv_resid<N, FMT>), dim3(grid_for((long long)m * (N / SGR) * SP_TS, 1)),
                                       bl, 0, st, c, m, A, X, Y, nullptr, nullptr);
            } else if (pass == 2 && FMT == SP_CSR) {
                hipLaunchKernelGGL((k_spmv_resid<N, FMT>), dim3(grid_for((long long)m * (N / SGR) * SP_TS, 1)), bl,
                                   0, st, c, m, A, X, Y, (const int *)perm, nlong);
            } else if (pass == 3 && FMT == SP_CSR) {
                hipLaunchKernelGGL((k_spmv_resid_wave<N>), dim3(rows_blocks(m)), dim3(BLK), 0, st, c, A, X, Y,
                                   (const int *)perm, nlong);
            } else if (pass == 1) {
                const unsigned g1 = grid_for((long long)m * tpr, L);
                hipLaunchKernelGGL((k_spmv<N, L, true, true, FMT>), dim3(g1 < 2048 ? g1 : 2048), bl, 0, st, c, m, A,
                                   X, Y, tpr, (const int *)rcnt, (const int *)rlist);
            }
        } else if (pass == 0) {
            hipLaunchKernelGGL((k_spmv<N, L, true, false, FMT>), gr, bl, 0, st, c, m, A, X, Y, tpr, nullptr, nullptr);
        }
    }
    static bool spmv_two(int fmt) { return SPLIT && fmt == SP_CSR; }
    static hipError_t spmv(const DevConsts *c, int pass, int fmt, int m, SpMat A, const char *x, char *y, int tpr,
                           int *ws, hipStream_t st) {
        if (m <= 0) return hipSuccess;
        if (tpr > BLK / L) return hipErrorInvalidValue;
        if (SPLIT && !ws) return hipErrorInvalidValue;
        switch (fmt) {
            case SP_CSR: spmv_fmt<SP_CSR>(c, pass, m, A, x, y, tpr, ws, st); break;
            case SP_ELL: spmv_fmt<SP_ELL>(c, pass, m, A, x, y, tpr, ws, st); break;
            case SP_JAD: spmv_fmt<SP_JAD>(c, pass, m, A, x, y, tpr, ws, st); break;
            default: spmv_fmt<SP_DIA>(c, pass, m, A, x, y, tpr, ws, st); break;
        }
        return hipGetLastError();
    }
    // ---- batched form (kernels_batch.hpp): clear the queue counter, then the persistent kernel ----
    static hipError_t batch(const DevConsts *c, const BatchItem *items, int count, int *queue, char *ws, int grid,
                            hipStream_t st) {
        if (count <= 0) return hipSuccess;
        hipLaunchKernelGGL((k_fill_int<N, L>), dim3(1), dim3(BLK), 0, st, queue, 1, 0);
        hipLaunchKernelGGL((k_batch<N, L>), dim3(grid), dim3(BLK), 0, st, c, items, count, queue, ws);
        return hipGetLastError();
    }
    // ---- the rest of the BLAS surface (kernels_blas.hpp) ----
    static hipError_t v1map(const DevConsts *c, int op, int sf, int m, int n, SoA S, int incs, SoA X, int ldx,
                            int incx, SoA Y, int ldy, int incy, SoA W, int ldw, int incw, hipStream_t st) {
        if ((long long)m * n <= 0) return hipSuccess;
        const dim3 gr(grid_for((long long)m * n, L)), bl(BLK);
#define MPH_V1(OP, SF)                                                                                            \
    hipLaunchKernelGGL((k_v1map<N, L, OP, SF>), gr, bl, 0, st, c, m, n, S, incs, X, ldx, incx, Y, ldy, incy, W, ldw, \
                       incw)
        if (op == V1_ADD)
            MPH_V1(V1_ADD, SF_SCALAR);
        else if (op == V1_SUB)
            MPH_V1(V1_SUB, SF_SCALAR);
        else if (sf == SF_ROW)
            MPH_V1(V1_MUL, SF_ROW);
        else if (sf == SF_COL)
            MPH_V1(V1_MUL, SF_COL);
        else
            MPH_V1(V1_MUL, SF_SCALAR);
#undef MPH_V1
        return hipGetLastError();
    }
    // sum of |x_i| (pass 0: the strided fold; the tree and second pass are fold / tree)
    static hipError_t fold_abs(const DevConsts *c, long long cnt, long long nvt, SoA x, int incx, char *out,
                               hipStream_t st) {
        if (nvt <= 0) return hipSuccess;
        SoA none{};
        hipLaunchKernelGGL((k_fold<N, L, 3, true, false>), dim3(grid_for(nvt, L)), dim3(BLK), 0, st, c, cnt, nvt, x,
                           incx, none, 1, AoSRef<N>{nullptr}, AoSRef<N>{nullptr}, AoSRef<N>{out});
        return hipGetLastError();
    }
    // maximum |x_i|: mode 0 the SoA vector (increment incx), 1 AoS partials; then the tree
    static hipError_t maxabs(const DevConsts *c, int mode, long long cnt, long long nvt, SoA x, int incx,
                             const char *xa, char *part, int b, unsigned p2, int blocks, char *out, hipStream_t st) {
        if (nvt <= 0) return hipSuccess;
        const dim3 gr(grid_for(nvt, L)), bl(BLK);
        if (mode == 0)
            hipLaunchKernelGGL((k_maxabs_fold<N, L, 0>), gr, bl, 0, st, c, cnt, nvt, x, incx, AoSRef<N>{(char *)xa},
                               AoSRef<N>{part});
        else
            hipLaunchKernelGGL((k_maxabs_fold<N, L, 1>), gr, bl, 0, st, c, cnt, nvt, x, incx, AoSRef<N>{(char *)xa},
                               AoSRef<N>{part});
        hipLaunchKernelGGL((k_maxabs_tree<N, L>), dim3(blocks), bl, 0, st, c, AoSRef<N>{part}, b, p2, AoSRef<N>{out});
        return hipGetLastError();
    }
    // per-row (cols false) or per-column sums of |A| over b1 virtual threads each
    static hipError_t matsum(const DevConsts *c, bool cols, int m, int n, SoA A, int lda, int b1, char *part,
                             hipStream_t st) {
        const long long nv = (long long)(cols ? n : m) * b1;
        if (nv <= 0) return hipSuccess;
        const dim3 gr(grid_for(nv, L)), bl(BLK);
        if (cols)
            hipLaunchKernelGGL((k_matsum_fold<N, L, true>), gr, bl, 0, st, c, m, n, A, lda, b1, AoSRef<N>{part});
        else
            hipLaunchKernelGGL((k_matsum_fold<N, L, false>), gr, bl, 0, st, c, m, n, A, lda, b1, AoSRef<N>{part});
        return hipGetLastError();
    }
    static hipError_t v2map(const DevConsts *c, int op, int m, int n, int upper, const char *alpha, const char *x,
                            int incx, const char *y, int incy, char *A, int lda, char *r, hipStream_t st) {
        if ((long long)m * n <= 0) return hipSuccess;
        const dim3 gr(grid_for((long long)m * n, L)), bl(BLK);
        const AoSRef<N> al{(char *)alpha}, X{(char *)x}, Y{(char *)y}, AA{A}, R{r};
#define MPH_V2(OP) hipLaunchKernelGGL((k_v2map<N, L, OP>), gr, bl, 0, st, c, m, n, upper, al, X, incx, Y, incy, AA, lda, R)
        switch (op) {
            case V2_SCAL: MPH_V2(V2_SCAL); break;
            case V2_GER: MPH_V2(V2_GER); break;
            case V2_SYR: MPH_V2(V2_SYR); break;
            default: MPH_V2(V2_SYR2); break;
        }
#undef MPH_V2
        return hipGetLastError();
    }
    static hipError_t symv_v2(const DevConsts *c, int upper, int n, const char *alpha, const char *A, int lda,
                              const char *x, int incx, const char *beta, char *y, int incy, hipStream_t st) {
        if (n <= 0) return hipSuccess;
        hipLaunchKernelGGL((k_symv_v2<N, L>), dim3(grid_for(n, L)), dim3(BLK), 0, st, c, upper, n,
                           AoSRef<N>{(char *)alpha}, AoSRef<N>{(char *)A}, lda, AoSRef<N>{(char *)x}, incx,
                           AoSRef<N>{(char *)beta}, AoSRef<N>{y}, incy);
        return hipGetLastError();
    }
    static hipError_t syrk_v2(const DevConsts *c, int upper, int trans, int n, int k, const char *alpha, const char *A,
                              int lda, const char *beta, char *Cm, int ldc, hipStream_t st) {
        if (n <= 0) return hipSuccess;
        hipLaunchKernelGGL((k_syrk_v2<N, L>), dim3(grid_for((long long)n * n, L)), dim3(BLK), 0, st, c, upper, trans,
                           n, k, AoSRef<N>{(char *)alpha}, AoSRef<N>{(char *)A}, lda, AoSRef<N>{(char *)beta},
                           AoSRef<N>{Cm}, ldc);
        return hipGetLastError();
    }
    static Ops table() {
        Ops o;
        o.N = N;
        o.L = L;
        o.set_d = &set_d;
        o.elem = &elem;
        o.cmp = &cmp;
        o.exp_out = &exp_out;
        o.fold = &fold;
        o.tree = &tree;
        o.aos_to_soa = &aos_to_soa;
        o.soa_to_aos = &soa_to_aos;
        o.axpy = &axpy;
        o.axpy_v2 = &axpy_v2;
        o.hostop = &hostop;
        o.v1map = &v1map;
        o.fold_abs = &fold_abs;
        o.maxabs = &maxabs;
        o.matsum = &matsum;
        o.v2map = &v2map;
        o.symv_v2 = &symv_v2;
        o.syrk_v2 = &syrk_v2;
        o.scal = &scal;
        o.gemv_fold = &gemv_fold;
        o.gemv_split = &gemv_split;
        o.gemv_final = &gemv_final;
        o.gemm = &gemm;
        o.gemm_ws = &gemm_ws;
        o.gemv_aux = &gemv_aux;
        o.gemm_mfma = &gemm_mfma;
        o.gemm_mfma_ws = &gemm_mfma_ws;
        o.gemm_v2 = &gemm_v2;
        o.spmv = &spmv;
        o.spmv_two = &spmv_two;
        o.batch = &batch;
        return o;
    }
};

}  // namespace mph

#define MPH_INSTANTIATE(N, L)                                \
    namespace mph {                                          \
    const Ops *ops_n##N##_l##L() {                           \
        static const Ops t = Launch<N, L>::table();          \
        return &t;                                           \
    }                                                        \
    }
#endif
