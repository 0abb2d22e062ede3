/*
 * Oracle: instantiation of orc_impl.h for n in {8,16,32} x policy {E,D}, and
 * the exported C entry points (mpres_oracle.h).
 * TEST INFRASTRUCTURE ONLY (see mpres_oracle.h).
 */
#include <fenv.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include <gmp.h>
#include <mpfr.h>

#include "orc_internal.h"

#define ORC_N 8
#define ORC_POL 0
#include "orc_impl.h"
#undef ORC_POL
#define ORC_POL 1
#include "orc_impl.h"
#undef ORC_POL
#undef ORC_N

#define ORC_N 16
#define ORC_POL 0
#include "orc_impl.h"
#undef ORC_POL
#define ORC_POL 1
#include "orc_impl.h"
#undef ORC_POL
#undef ORC_N

#define ORC_N 32
#define ORC_POL 0
#include "orc_impl.h"
#undef ORC_POL
#define ORC_POL 1
#include "orc_impl.h"
#undef ORC_POL
#undef ORC_N

void orc_set_threads(int nthreads) {
#ifdef _OPENMP
    omp_set_num_threads(nthreads > 0 ? nthreads : 1);
#else
    (void)nthreads;
#endif
}

/* Dispatch on (n, policy): DISPATCH(expr) expands expr with F(x) -> x_<pol><n> and MPF -> mpf_<pol><n>. */
#define DISPATCH(c, policy, BODY)                                               \
    do {                                                                        \
        if ((c)->n == 8 && (policy) == 0) { BODY(_e8) }                         \
        else if ((c)->n == 8) { BODY(_d8) }                                     \
        else if ((c)->n == 16 && (policy) == 0) { BODY(_e16) }                  \
        else if ((c)->n == 16) { BODY(_d16) }                                   \
        else if ((c)->n == 32 && (policy) == 0) { BODY(_e32) }                  \
        else { BODY(_d32) }                                                     \
    } while (0)
#define DISPATCH_D(c, BODY)                                                     \
    do {                                                                        \
        if ((c)->n == 8) { BODY(_d8) }                                          \
        else if ((c)->n == 16) { BODY(_d16) }                                   \
        else { BODY(_d32) }                                                     \
    } while (0)
#define G(x, s) x##s

/* ---- scalar primitives ---- */
double orc_dadd(int policy, int up, double a, double b) {
    if (policy == 0) return up ? dadd_ru_e8(a, b) : dadd_rd_e8(a, b);
    return up ? dadd_ru_d8(a, b) : dadd_rd_d8(a, b);
}
double orc_dsub(int policy, int up, double a, double b) {
    if (policy == 0) return up ? dsub_ru_e8(a, b) : dsub_rd_e8(a, b);
    return up ? dsub_ru_d8(a, b) : dsub_rd_d8(a, b);
}
double orc_dmul(int policy, int up, double a, double b) {
    if (policy == 0) return up ? dmul_ru_e8(a, b) : dmul_rd_e8(a, b);
    return up ? dmul_ru_d8(a, b) : dmul_rd_d8(a, b);
}
double orc_ddiv(int policy, int up, double a, double b) {
    if (policy == 0) return up ? ddiv_ru_e8(a, b) : ddiv_rd_e8(a, b);
    return up ? ddiv_ru_d8(a, b) : ddiv_rd_d8(a, b);
}
double orc_psum(int policy, int up, int n, const double *x) {
    if (n == 8) {
        if (policy == 0) return up ? psum_ru_e8(x) : psum_rd_e8(x);
        return up ? psum_ru_d8(x) : psum_rd_d8(x);
    } else if (n == 16) {
        if (policy == 0) return up ? psum_ru_e16(x) : psum_rd_e16(x);
        return up ? psum_ru_d16(x) : psum_rd_d16(x);
    }
    if (policy == 0) return up ? psum_ru_e32(x) : psum_rd_e32(x);
    return up ? psum_ru_d32(x) : psum_rd_d32(x);
}
int orc_ceil_log2_lib(double x) { return (int)ceil(log2(x)); }

/* ---- element operations ---- */
#define B_SET_D(s)                                                                         \
    _Pragma("omp parallel for schedule(static)")                                           \
    for (int i = 0; i < count; i++) G(set_d, s)(c, &((G(mpf, s) *)out)[i], x[i]);
void orc_set_d(const orc_ctx *c, int policy, int count, const double *x, void *out) { DISPATCH(c, policy, B_SET_D); }

#define B_ADD(s)                                                                           \
    for (int i = 0; i < count; i++)                                                        \
        G(add, s)(c, &((G(mpf, s) *)r)[i], ((const G(mpf, s) *)x)[i], ((const G(mpf, s) *)y)[i]);
void orc_add(const orc_ctx *c, int policy, int count, const void *x, const void *y, void *r) { DISPATCH(c, policy, B_ADD); }

/* mp_sub: src/arith/sub.cuh:31-54 (sign of y flipped, then mp_add) */
#define B_SUB(s)                                                                           \
    for (int i = 0; i < count; i++) {                                                      \
        G(mpf, s) yy = ((const G(mpf, s) *)y)[i];                                          \
        yy.sign ^= 1;                                                                      \
        G(add, s)(c, &((G(mpf, s) *)r)[i], ((const G(mpf, s) *)x)[i], yy);                 \
    }
void orc_sub(const orc_ctx *c, int policy, int count, const void *x, const void *y, void *r) { DISPATCH(c, policy, B_SUB); }

#define B_MUL(s)                                                                           \
    for (int i = 0; i < count; i++)                                                        \
        G(mul, s)(c, &((G(mpf, s) *)r)[i], ((const G(mpf, s) *)x)[i], ((const G(mpf, s) *)y)[i]);
void orc_mul(const orc_ctx *c, int policy, int count, const void *x, const void *y, void *r) { DISPATCH(c, policy, B_MUL); }

#define B_MULD(s)                                                                          \
    for (int i = 0; i < count; i++) G(mul_d, s)(c, &((G(mpf, s) *)r)[i], ((const G(mpf, s) *)x)[i], y[i]);
void orc_mul_d(const orc_ctx *c, int policy, int count, const void *x, const double *y, void *r) {
    DISPATCH(c, policy, B_MULD);
}

#define B_VADD(s)                                                                          \
    for (int i = 0; i < count; i++)                                                        \
        G(vadd, s)(c, &((G(mpf, s) *)r)[i], ((const G(mpf, s) *)x)[i], ((const G(mpf, s) *)y)[i]);
void orc_vadd(const orc_ctx *c, int count, const void *x, const void *y, void *r) { DISPATCH_D(c, B_VADD); }

#define B_ROUND(s)                                                                         \
    for (int i = 0; i < count; i++) G(round_if_needed, s)(c, &((G(mpf, s) *)x)[i]);
void orc_round(const orc_ctx *c, int policy, int count, void *x) { DISPATCH(c, policy, B_ROUND); }

#define B_SCALE(s)                                                                         \
    for (int i = 0; i < count; i++) G(scale2pow, s)(c, r + (long)i * c->n, x + (long)i * c->n, D[i]);
void orc_scale2pow(const orc_ctx *c, int policy, int count, const int *x, const unsigned *D, int *r) {
    DISPATCH(c, policy, B_SCALE);
}

#define B_EVAL(s)                                                                          \
    for (int i = 0; i < count; i++) G(eval_compute, s)(c, &low[i], &upp[i], x + (long)i * c->n, fast);
void orc_eval(const orc_ctx *c, int policy, int fast, int count, const int *x, orc_er_t *low, orc_er_t *upp) {
    DISPATCH(c, policy, B_EVAL);
}

#define B_RANK(s)                                                                          \
    for (int i = 0; i < count; i++) {                                                      \
        int sv[32];                                                                        \
        const int *xi = x + (long)i * c->n;                                                \
        G(rns_mul, s)(c, sv, xi, c->w);                                                    \
        k[i] = fast ? G(rank_fast, s)(c, sv) : G(rank_full, s)(c, xi, sv);                 \
    }
void orc_rank(const orc_ctx *c, int policy, int fast, int count, const int *x, int *k) { DISPATCH(c, policy, B_RANK); }

#define B_MRD(s)                                                                           \
    for (int i = 0; i < count; i++) mrd[i] = G(mrd, s)(c, x + (long)i * c->n);
void orc_mrd(const orc_ctx *c, int count, const int *x, int *mrd) { DISPATCH_D(c, B_MRD); }

/* rns_cmp, src/rns.cuh:1176-1225 (CUDA :1242-1291) */
#define B_CMP(s)                                                                           \
    for (int i = 0; i < count; i++) {                                                      \
        const int *xi = x + (long)i * c->n, *yi = y + (long)i * c->n;                      \
        orc_er_t xl, xu, yl, yu;                                                           \
        G(eval_compute, s)(c, &xl, &xu, xi, 0);                                            \
        G(eval_compute, s)(c, &yl, &yu, yi, 0);                                            \
        if (G(er_ucmp, s)(xl, yu) > 0) { cmp[i] = 1; continue; }                           \
        if (G(er_ucmp, s)(yl, xu) > 0) { cmp[i] = -1; continue; }                          \
        int eq = 1;                                                                        \
        for (int q = 0; q < c->n; q++) if (xi[q] != yi[q]) { eq = 0; break; }              \
        cmp[i] = eq ? 0 : G(mrc_compare, s)(c, xi, yi);                                    \
    }
void orc_rns_cmp(const orc_ctx *c, int policy, int count, const int *x, const int *y, int *cmp) {
    DISPATCH(c, policy, B_CMP);
}

/* mp_cmp / mp_cmp_abs, src/arith/cmp.cuh, cmpabs.cuh */
#define B_MPCMP(s)                                                                         \
    for (int i = 0; i < count; i++)                                                        \
        r[i] = G(cmp, s)(c, ((const G(mpf, s) *)x)[i], ((const G(mpf, s) *)y)[i], abs_only);
void orc_mp_cmp(const orc_ctx *c, int policy, int abs_only, int count, const void *x, const void *y, int *r) {
    DISPATCH(c, policy, B_MPCMP);
}

/* ---- compositions ---- */
#define B_DOT(s)                                                                           \
    G(dot, s)(c, n, (const G(mpf, s) *)x, incx, (const G(mpf, s) *)y, incy, g3, b3, (G(mpf, s) *)r);
void orc_dot(const orc_ctx *c, int policy, int n, const void *x, int incx, const void *y, int incy, int g3, int b3,
             void *r) {
    DISPATCH(c, policy, B_DOT);
}

#define B_AXPY(s)                                                                          \
    G(axpy, s)(c, n, (const G(mpf, s) *)alpha, (const G(mpf, s) *)x, incx, (G(mpf, s) *)y, incy);
void orc_axpy(const orc_ctx *c, int policy, int n, const void *alpha, const void *x, int incx, void *y, int incy) {
    DISPATCH(c, policy, B_AXPY);
}

#define B_AXPY2(s)                                                                         \
    G(axpy_v2, s)(c, n, (const G(mpf, s) *)alpha, (const G(mpf, s) *)x, (const G(mpf, s) *)y, (G(mpf, s) *)r);
void orc_axpy_v2(const orc_ctx *c, int policy, int n, const void *alpha, const void *x, const void *y, void *r) {
    DISPATCH(c, policy, B_AXPY2);
}

#define B_GEMV(s)                                                                          \
    G(gemv, s)(c, trans, m, n, (const G(mpf, s) *)alpha, (const G(mpf, s) *)A, lda, (const G(mpf, s) *)x, incx, \
               (const G(mpf, s) *)beta, (G(mpf, s) *)y, incy, b3, rows, nrows);
void orc_gemv(const orc_ctx *c, int policy, int trans, int m, int n, const void *alpha, const void *A, int lda,
              const void *x, int incx, const void *beta, void *y, int incy, int b3, const int *rows, int nrows) {
    DISPATCH(c, policy, B_GEMV);
}

#define B_GEMM(s)                                                                          \
    G(gemm, s)(c, m, n, k, (const G(mpf, s) *)alpha, (const G(mpf, s) *)A, lda, (const G(mpf, s) *)B, ldb, \
               (const G(mpf, s) *)beta, (G(mpf, s) *)C, ldc);
void orc_gemm(const orc_ctx *c, int policy, int m, int n, int k, const void *alpha, const void *A, int lda,
              const void *B, int ldb, const void *beta, void *C, int ldc) {
    DISPATCH(c, policy, B_GEMM);
}

#define B_GEMM2(s)                                                                         \
    G(gemm_v2, s)(c, transa, transb, m, n, k, (const G(mpf, s) *)alpha, (const G(mpf, s) *)A, lda,            \
                  (const G(mpf, s) *)B, ldb, (const G(mpf, s) *)beta, (G(mpf, s) *)C, ldc, entries, nentries);
void orc_gemm_v2(const orc_ctx *c, int policy, int transa, int transb, int m, int n, int k, const void *alpha,
                 const void *A, int lda, const void *B, int ldb, const void *beta, void *C, int ldc,
                 const int *entries, int nentries) {
    DISPATCH(c, policy, B_GEMM2);
}

#define B_SPMV(s)                                                                          \
    G(spmv, s)(c, m, irp, ja, as, (const G(mpf, s) *)x, (G(mpf, s) *)y, tpr, rows, nrows);
void orc_spmv_csr(const orc_ctx *c, int policy, int m, const int *irp, const int *ja, const double *as, const void *x,
                  void *y, int tpr, const int *rows, int nrows) {
    DISPATCH(c, policy, B_SPMV);
}

/* ---- v2 helpers of the solvers, mp_add_abs, sum of squares ---- */
#define B_ADDABS(s)                                                                        \
    for (int i = 0; i < count; i++)                                                        \
        G(add_abs, s)(c, &((G(mpf, s) *)r)[i], ((const G(mpf, s) *)x)[i], ((const G(mpf, s) *)y)[i]);
void orc_add_abs(const orc_ctx *c, int policy, int count, const void *x, const void *y, void *r) {
    DISPATCH(c, policy, B_ADDABS);
}
#define B_MAXPY(s) G(maxpy_v2, s)(c, n, (const G(mpf, s) *)alpha, (const G(mpf, s) *)x, (const G(mpf, s) *)y, (G(mpf, s) *)r);
void orc_maxpy_v2(const orc_ctx *c, int policy, int n, const void *alpha, const void *x, const void *y, void *r) {
    DISPATCH(c, policy, B_MAXPY);
}
#define B_SUMSQ(s) G(sumsq, s)(c, n, (const G(mpf, s) *)x, blocks, threads, (G(mpf, s) *)r);
void orc_sumsq(const orc_ctx *c, int policy, int n, const void *x, int blocks, int threads, void *r) {
    DISPATCH(c, policy, B_SUMSQ);
}

/* ---- host MPFR operations (policy-independent) ---- */
#define DISPATCH_E(c, BODY)                                                     \
    do {                                                                        \
        if ((c)->n == 8) { BODY(_e8) }                                          \
        else if ((c)->n == 16) { BODY(_e16) }                                   \
        else { BODY(_e32) }                                                     \
    } while (0)
#define B_DIV(s)                                                                           \
    for (int i = 0; i < count; i++)                                                        \
        G(div, s)(c, &((G(mpf, s) *)r)[i], &((const G(mpf, s) *)x)[i], &((const G(mpf, s) *)y)[i]);
void orc_div(const orc_ctx *c, int count, const void *x, const void *y, void *r) { DISPATCH_E(c, B_DIV); }
#define B_SQRT(s)                                                                          \
    for (int i = 0; i < count; i++) G(sqrt, s)(c, &((G(mpf, s) *)r)[i], &((const G(mpf, s) *)x)[i]);
void orc_sqrt(const orc_ctx *c, int count, const void *x, void *r) { DISPATCH_E(c, B_SQRT); }
#define B_DSQRT(s)                                                                         \
    for (int i = 0; i < count; i++) out[i] = G(dsqrt, s)(c, &((const G(mpf, s) *)x)[i]);
void orc_dsqrt(const orc_ctx *c, int count, const void *x, double *out) { DISPATCH_E(c, B_DSQRT); }

/* ELL / JAD / DIA SpMV */
#define B_SPFMT(s) G(spmv_fmt, s)(c, fmt, m, w, ncols, as, ja, jcp, perm, offset, (const G(mpf, s) *)x, (G(mpf, s) *)y, tpr);
void orc_spmv_fmt(const orc_ctx *c, int policy, int fmt, int m, int w, int ncols, const double *as, const int *ja,
                  const int *jcp, const int *perm, const int *offset, const void *x, void *y, int tpr) {
    DISPATCH(c, policy, B_SPFMT);
}
