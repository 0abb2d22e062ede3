/*
 * CPU oracle for the RNS multiple-precision hot path of MPRES-BLAS.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * never by the product library (mpres-blas_amd/), which must fail loudly when
 * its HIP code is missing instead of falling back to anything here.
 *
 * It is a plain-C restatement of the reference's scalar algorithms under two
 * rounding policies, because the reference's CPU and CUDA code paths differ:
 *
 *   ORC_POLICY_E (0): the reference CPU path. Directed rounding is emulated as
 *       RN(op) -/+ (phi*|c| + eta)  (src/dinterval.cuh:55-165), moduli
 *       products use a double-reciprocal quotient (src/modular.cuh:90-95),
 *       pairwise sums bump at the root (src/pairwise.cuh:402-479).
 *   ORC_POLICY_D (1): the reference CUDA path. True IEEE-754 directed rounding
 *       (__dadd_rd & co., here via fesetround), exact '%' reductions
 *       (src/modular.cuh:118-314), src/rns.cuh:797-1160 and the cuda::
 *       overloads in src/arith/.
 *
 * The HIP kernels are required to match ORC_POLICY_D bit for bit in every
 * field (residues, sign, exponent and both interval bounds).
 *
 * Numbers cross this interface in the reference's mp_float_t layout
 * (src/types.cuh:69-74): int digits[n]; int sign; int exp; er_float_t eval[2],
 * i.e. 4n+40 bytes per element, n in {8,16,32}.
 *
 * Parity pinning of each function is listed in the header comment of
 * oracle/orc_impl.h (which checks cover which function).
 */
#ifndef MPRES_ORACLE_H
#define MPRES_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_POLICY_E 0
#define ORC_POLICY_D 1

/* er_float_t, src/types.cuh:46-49 */
typedef struct { double frac; long exp; } orc_er_t;

typedef struct orc_ctx orc_ctx;

/* Moduli-set identifiers (one per reference params file with n in {8,16,32}). */
int orc_num_sets(void);
const char *orc_set_name(int set_id);

orc_ctx *orc_ctx_create(int set_id);
void orc_ctx_destroy(orc_ctx *c);
void orc_set_threads(int nthreads);

/* ---- constants (for the closed-form tests) ---- */
int orc_n(const orc_ctx *c);
int orc_log2m(const orc_ctx *c);                 /* RNS_MODULI_PRODUCT_LOG2 of the params file */
void orc_moduli(const orc_ctx *c, int *out);     /* n */
void orc_part_inverse(const orc_ctx *c, int *out);/* n: (M/m_i)^-1 mod m_i */
void orc_pow2_table(const orc_ctx *c, int *out); /* (L+1) x n */
void orc_pow2_residues(const orc_ctx *c, int *out);       /* 30: M mod 2^(j+1) */
void orc_part_pow2_residues(const orc_ctx *c, int *out);  /* 30 x n: M_i mod 2^(j+1) */
void orc_pow2_inverse(const orc_ctx *c, int *out);        /* 30 x n: 2^-(j+1) mod m_i */
void orc_mrc_inverse(const orc_ctx *c, int *out);         /* n x n */
void orc_recip(const orc_ctx *c, double *rd, double *ru); /* n each */
void orc_eval_consts(const orc_ctx *c, double *acc, int *ref_factor,
                     orc_er_t *unit_low, orc_er_t *unit_upp,
                     orc_er_t *inv_unit_low, orc_er_t *inv_unit_upp);
void orc_mp_consts(const orc_ctx *c, int *precision, int *h, int *j);

/* ---- directed-rounding primitives (policy-selected), for pinning ---- */
double orc_dadd(int policy, int up, double a, double b);
double orc_dsub(int policy, int up, double a, double b);
double orc_dmul(int policy, int up, double a, double b);
double orc_ddiv(int policy, int up, double a, double b);
double orc_psum(int policy, int up, int n, const double *x);
int orc_ceil_log2_lib(double x);   /* ceil(log2(x)) exactly as the reference spells it */

/* ---- element operations on AoS mp_float_t arrays (count elements) ---- */
void orc_set_d(const orc_ctx *c, int policy, int count, const double *x, void *out);
void orc_add(const orc_ctx *c, int policy, int count, const void *x, const void *y, void *r);
void orc_sub(const orc_ctx *c, int policy, int count, const void *x, const void *y, void *r);
void orc_mul(const orc_ctx *c, int policy, int count, const void *x, const void *y, void *r);
void orc_mul_d(const orc_ctx *c, int policy, int count, const void *x, const double *y, void *r);
/* v1 vector-add esi+digits+round kernels (src/mpvector.cuh:234-472,669-714): no MRC, quotient digits */
void orc_vadd(const orc_ctx *c, int count, const void *x, const void *y, void *r);
void orc_round(const orc_ctx *c, int policy, int count, void *x);
/* digits-only helpers: x is count*n ints */
void orc_scale2pow(const orc_ctx *c, int policy, int count, const int *x, const unsigned *D, int *r);
void orc_eval(const orc_ctx *c, int policy, int fast, int count, const int *x, orc_er_t *low, orc_er_t *upp);
void orc_rank(const orc_ctx *c, int policy, int fast, int count, const int *x, int *k);
void orc_mrd(const orc_ctx *c, int count, const int *x, int *mrd);
void orc_rns_cmp(const orc_ctx *c, int policy, int count, const int *x, const int *y, int *cmp);
/* mp_cmp (abs_only = 0) / mp_cmp_abs (abs_only = 1) of AoS numbers: r[i] in {-1, 0, 1} */
void orc_mp_cmp(const orc_ctx *c, int policy, int abs_only, int count, const void *x, const void *y, int *r);

/* ---- BLAS compositions in the reference's operation order ---- */
/* dot: products then the two-pass reduction for launch shape (g3,b3), src/blas/dot.cuh:41-107;
 * identical to v2 mp_dot<g3,b3> (src/blas/v2/dot_v2.cuh:36-129). */
void orc_dot(const orc_ctx *c, int policy, int n, const void *x, int incx, const void *y, int incy,
             int g3, int b3, void *r);
/* v1 axpy, src/blas/axpy.cuh:45-74 (y updated in place) */
void orc_axpy(const orc_ctx *c, int policy, int n, const void *alpha, const void *x, int incx, void *y, int incy);
/* v2 axpy, src/blas/v2/axpy_v2.cuh:39-52 */
void orc_axpy_v2(const orc_ctx *c, int policy, int n, const void *alpha, const void *x, const void *y, void *r);
/* v1 gemv, src/blas/gemv.cuh:150-268, trans: 0 = no-trans, 1 = trans; column-major */
void orc_gemv(const orc_ctx *c, int policy, int trans, int m, int n, const void *alpha, const void *A, int lda,
              const void *x, int incx, const void *beta, void *y, int incy, int b3, const int *rows, int nrows);
/* v1 gemm (no-trans only), src/blas/gemm.cuh:39-167 */
void orc_gemm(const orc_ctx *c, int policy, int m, int n, int k, const void *alpha, const void *A, int lda,
              const void *B, int ldb, const void *beta, void *C, int ldc);
/* v2 gemm kernel, src/blas/v2/gemm_v2.cuh:55-81; transa/transb: 0 = N, 1 = T */
void orc_gemm_v2(const orc_ctx *c, int policy, int transa, int transb, int m, int n, int k, const void *alpha,
                 const void *A, int lda, const void *B, int ldb, const void *beta, void *C, int ldc,
                 const int *entries, int nentries);
/* SpMV CSR scalar (tpr = 1, src/sparse/spmv/spmv_csr.cuh:48-64) and vector (tpr > 1, spmv_csrv.cuh:48-87) */
void orc_spmv_csr(const orc_ctx *c, int policy, int m, const int *irp, const int *ja, const double *as,
                  const void *x, void *y, int tpr, const int *rows, int nrows);

#ifdef __cplusplus
}
#endif
#endif
