/*
 * mpres-hip: MI355X (gfx950) implementation of the RNS multiple-precision hot
 * path of MPRES-BLAS, as a C ABI.
 *
 * Data types are layout-identical to the reference (src/types.cuh:46-113), so
 * buffers built for the reference can be handed over unchanged:
 *   er_float_t       {double frac; long exp}                       16 bytes
 *   mp_float_t<n>    {int digits[n]; int sign; int exp; er_float_t eval[2]}  4n+40 bytes
 *   mp_array_t       structure of arrays with buf and a device-side len
 *   mp_collection_t  structure of arrays without buf/len
 *   csr_t            {double *as; int *ja; int *irp}
 * All pointers inside these structs are device pointers. The number of moduli
 * n is a property of the context (one of the reference params files), not of
 * the type, so mp_float_t arrays are passed as void pointers with a stride of
 * mpres_mp_float_size(ctx) bytes.
 *
 * Every entry point returns an int status (MPRES_OK or a negative error) and
 * enqueues its kernels on the caller's HIP stream (NULL = default stream).
 * Results are bit-identical to the reference's CUDA algorithm on the same
 * inputs and summation order: operations whose result depends on the order of
 * multiple-precision additions take the reference's launch parameters that
 * fix that order (the "order" arguments below) and reproduce it.
 *
 * Workspace: entry points that need scratch space take (ws, ws_bytes). Pass
 * ws == NULL to use a buffer owned by the context (allocated on first use and
 * grown when needed); the matching mpres_*_workspace() call returns the size.
 */
#ifndef MPRES_HIP_H
#define MPRES_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define MPRES_OK 0
#define MPRES_ERR_ARG (-1)        /* invalid argument (size, stride, leading dimension, order) */
#define MPRES_ERR_HIP (-2)        /* a HIP runtime call failed; see mpres_last_error() */
#define MPRES_ERR_WORKSPACE (-3)  /* caller workspace too small */
#define MPRES_ERR_UNSUPPORTED (-4)/* valid call the library does not implement (e.g. transposed v1 gemm) */
#define MPRES_ERR_NO_DEVICE (-5)  /* no gfx950 device */

/* ---- reference types (src/types.cuh) ---- */
typedef struct { double frac; long exp; } er_float_t;                 /* src/types.cuh:46-49 */
typedef struct { int x, y, z, w; } mpres_int4_t;                      /* CUDA int4 */
typedef struct {                                                      /* src/types.cuh:85-92 */
    int *digits;
    int *sign;
    int *exp;
    er_float_t *eval;
    mpres_int4_t *buf;
    int *len;
} mp_array_t;
typedef struct {                                                      /* src/types.cuh:99-104 */
    int *digits;
    int *sign;
    int *exp;
    er_float_t *eval;
} mp_collection_t;
typedef struct { double *as; int *ja; int *irp; } csr_t;              /* src/types.cuh:109-113 */
typedef struct { double *as; int *ja; int *jcp; int *perm; } jad_t;   /* src/types.cuh:118-123 */
typedef struct { double *as; int *ja; } ell_t;                        /* src/types.cuh:128-131 */
typedef struct { double *as; int *offset; } dia_t;                    /* src/types.cuh:136-139 */

/* The layouts above are the reference's (LP64: long and pointers are 8 bytes). */
#ifdef __cplusplus
#define MPRES_STATIC_ASSERT(c, msg) static_assert(c, msg)
#else
#define MPRES_STATIC_ASSERT(c, msg) _Static_assert(c, msg)
#endif
MPRES_STATIC_ASSERT(sizeof(long) == 8 && sizeof(void *) == 8, "LP64 data model required");
MPRES_STATIC_ASSERT(sizeof(er_float_t) == 16 && offsetof(er_float_t, exp) == 8, "er_float_t layout");
MPRES_STATIC_ASSERT(sizeof(mpres_int4_t) == 16, "int4 layout");
MPRES_STATIC_ASSERT(sizeof(mp_array_t) == 48 && offsetof(mp_array_t, sign) == 8 && offsetof(mp_array_t, exp) == 16 &&
                        offsetof(mp_array_t, eval) == 24 && offsetof(mp_array_t, buf) == 32 &&
                        offsetof(mp_array_t, len) == 40,
                    "mp_array_t layout");
MPRES_STATIC_ASSERT(sizeof(mp_collection_t) == 32 && offsetof(mp_collection_t, eval) == 24, "mp_collection_t layout");
MPRES_STATIC_ASSERT(sizeof(csr_t) == 24 && offsetof(csr_t, ja) == 8 && offsetof(csr_t, irp) == 16, "csr_t layout");
MPRES_STATIC_ASSERT(sizeof(jad_t) == 32 && offsetof(jad_t, jcp) == 16 && offsetof(jad_t, perm) == 24, "jad_t layout");
MPRES_STATIC_ASSERT(sizeof(ell_t) == 16 && offsetof(ell_t, ja) == 8, "ell_t layout");
MPRES_STATIC_ASSERT(sizeof(dia_t) == 16 && offsetof(dia_t, offset) == 8, "dia_t layout");

/* mblas_trans_type, src/blas/mblas_enum.cuh:25-31 */
#define MPRES_NO_TRANS 111
#define MPRES_TRANS 112
#define MPRES_CONJ_TRANS 113
/* mblas_uplo_type, mblas_side_type, mblas_norm_type: src/blas/mblas_enum.cuh:31-44 */
#define MPRES_UPPER 121
#define MPRES_LOWER 122
#define MPRES_LEFT 141
#define MPRES_RIGHT 142
#define MPRES_ONE_NORM 171
#define MPRES_INF_NORM 175

typedef struct mpres_ctx mpres_ctx_t;

/* ---- library / context ----
 * Replaces rns_const_init() + mp_const_init() (src/rns.cuh:324-442,
 * src/arith/arith_utils.cuh:44-85): the constants of one moduli set are
 * computed on the host and placed in device memory owned by the context. */
int mpres_num_moduli_sets(void);
const char *mpres_moduli_set_name(int set_id);   /* "8_2double", "16_4double", "32_8double", ... */
int mpres_moduli_set_id(const char *name);        /* -1 if unknown */
int mpres_ctx_create(mpres_ctx_t **ctx, int set_id, int device);
int mpres_ctx_destroy(mpres_ctx_t *ctx);
/* n, MP_PRECISION, MP_H, MP_J, RNS_MODULI_PRODUCT_LOG2 */
int mpres_ctx_info(const mpres_ctx_t *ctx, int *n, int *precision, int *mp_h, int *mp_j, int *log2m);
size_t mpres_mp_float_size(const mpres_ctx_t *ctx);  /* 4n + 40 */
const char *mpres_last_error(void);                  /* thread-local message of the last failure */
/* copies of the device constant tables, for checking against an independent recomputation:
 * what = 0 moduli (n ints), 1 (M/m_i)^-1 mod m_i (n), 2 RNS_POW2 ((log2m+1)*n), 3 M mod 2^(j+1) (30),
 * 4 M_i mod 2^(j+1) (30*n), 5 2^-(j+1) mod m_i (30*n), 6 MRC inverses (n*n);
 * what = 7 doubles: recip_rd (n), recip_ru (n), acc; what = 8 er_float_t: unit_low, unit_upp, inv_unit_low, inv_unit_upp;
 * what = 9 ints: MP_PRECISION, MP_H, MP_J, RNS_EVAL_REF_FACTOR, RNS_MODULI_PRODUCT_LOG2 */
int mpres_ctx_table(const mpres_ctx_t *ctx, int what, void *out, size_t bytes);
/* the same tables computed on the host for a moduli set, without a device */
int mpres_moduli_set_table(int set_id, int what, void *out, size_t bytes);

/* ---- mp_array_t management, src/mparray.cuh:35-165 ----
 * host2device / device2host take host arrays of mp_float_t (AoS, 4n+40 bytes each)
 * and move them in one transfer per field instead of one per element. */
int mpres_array_init(mpres_ctx_t *ctx, mp_array_t *a, size_t size);
int mpres_array_clear(mpres_ctx_t *ctx, mp_array_t *a);
int mpres_array_host2device(mpres_ctx_t *ctx, mp_array_t *dst, const void *host_src, size_t size, void *stream);
int mpres_array_device2host(mpres_ctx_t *ctx, void *host_dst, const mp_array_t *src, size_t size, void *stream);

/* ---- double -> mp conversion on the GPU: mp_set_d, src/arith/assign.cuh:54-81 ----
 * (the reference converts on the CPU only). x and dst are device pointers. */
int mpres_array_set_d(mpres_ctx_t *ctx, mp_array_t *dst, const double *x, int count, void *stream);
int mpres_set_d(mpres_ctx_t *ctx, void *dst_aos, const double *x, int count, void *stream);
/* AoS <-> SoA on the device (mp_float_t[] <-> mp_array_t, first count elements) */
int mpres_aos_to_array(mpres_ctx_t *ctx, mp_array_t *dst, const void *src_aos, int count, void *stream);
int mpres_array_to_aos(mpres_ctx_t *ctx, void *dst_aos, const mp_array_t *src, int count, void *stream);

/* ---- element-wise arithmetic on AoS arrays (cuda::mp_add / mp_sub / mp_mul / mp_mul_d,
 * src/arith/{add,sub,mul,muld}.cuh); r[i] = x[i] op y[i] ---- */
#define MPRES_OP_ADD 0
#define MPRES_OP_SUB 1
#define MPRES_OP_MUL 2
#define MPRES_OP_MUL_D 3   /* y is a device array of double */
#define MPRES_OP_VADD 4    /* the v1 vector-add kernels: src/mpvector.cuh:234-472 + round :669-714 */
int mpres_elementwise(mpres_ctx_t *ctx, int op, int count, const void *x, const void *y, void *r, void *stream);

/* ---- comparison ----
 * r[i] = 1, 0 or -1 as x[i] >, = or < y[i] (MPRES_CMP: cuda::mp_cmp, src/arith/cmp.cuh:79-130)
 * or as |x[i]| >, = or < |y[i]| (MPRES_CMP_ABS: cuda::mp_cmp_abs, src/arith/cmpabs.cuh:76-119).
 * x, y: device arrays of count mp_float_t; r: device array of count ints. The operands are
 * aligned to the larger exponent as the reference does (an operand below 2^MP_J of the other
 * counts as zero), then compared by interval evaluation, falling back to mixed-radix digits. */
#define MPRES_CMP 0
#define MPRES_CMP_ABS 1
int mpres_cmp(mpres_ctx_t *ctx, int mode, int count, const void *x, const void *y, int *r, void *stream);
/* r[i] = 1, 0 or -1 as the RNS integers x_i >, = or < y_i: cuda::rns_cmp(x, y),
 * src/rns.cuh:1242-1260. x, y: device arrays of count x n residues (0 <= x_i < M). */
int mpres_rns_cmp(mpres_ctx_t *ctx, int count, const int *x, const int *y, int *r, void *stream);

/* ---- the operations the reference runs on the host with MPFR, on the GPU ----
 * r[i] = x[i] / y[i]: cuda::mp_div (src/arith/div.cuh:56-65, a host MPFR round trip there):
 * both operands rounded to MP_PRECISION bits, the quotient rounded to nearest at
 * MP_PRECISION bits, stored as mp_set_mpfr stores it (significand without trailing zeros)
 * with the CPU interval evaluation of its residues. A zero divisor gives zero (undefined in
 * the reference). x, y, r: device arrays of count mp_float_t. */
int mpres_div(mpres_ctx_t *ctx, int count, const void *x, const void *y, void *r, void *stream);
/* cuda::mp_sqrt / cuda::mp_dsqrt (src/arith/sqrt.cuh:65-89): the square root rounded to
 * nearest at MP_PRECISION bits, as an mp_float_t or rounded to a double (out: device array).
 * A negative operand gives NaN (mp_dsqrt) or zero (mp_sqrt). */
int mpres_sqrt(mpres_ctx_t *ctx, int count, const void *x, void *r, void *stream);
int mpres_dsqrt(mpres_ctx_t *ctx, int count, const void *x, double *out, void *stream);

/* ---- export (replaces the host-side mp_get_d / mp_get_mpfr, src/arith/assign.cuh:154-211) ----
 * out[i] = x[i] as a double: (-1)^sign X 2^exp with X the RNS integer of the residues (CRT
 * mod M), rounded to 53 bits to nearest, ties to even, then to the double format (subnormal
 * results round once more, overflow gives an infinity) - the reference's MPFR conversion at
 * 53-bit precision. x: count mp_float_t (mpres_get_d) or an mp_array_t (mpres_array_get_d);
 * out: device array of count doubles. */
int mpres_get_d(mpres_ctx_t *ctx, int count, const void *x, double *out, void *stream);
int mpres_array_get_d(mpres_ctx_t *ctx, int count, const mp_array_t *x, double *out, void *stream);
/* The exact value of each number for MPFR or decimal output: limbs[i * L + l] = 32-bit limb l
 * (least significant first) of X, with L = mpres_export_limbs(ctx); sign[i], exp[i] as stored.
 * x[i] = (-1)^sign[i] X 2^exp[i]. With GMP/MPFR: mpz_import(z, L, -1, 4, 0, 0, limbs + i * L),
 * mpfr_set_z_2exp(r, z, exp[i], MPFR_RNDN), then negate if sign[i]. All arrays on the device. */
int mpres_export_limbs(const mpres_ctx_t *ctx);
int mpres_get_integer(mpres_ctx_t *ctx, int count, const void *x, unsigned *limbs, int *sign, int *exp,
                      void *stream);

/* ---- BLAS, v1 interface over mp_array_t ---- */
/* r = x . y ; cuda::mp_dot<gridDim1,blockDim1,gridDim2,gridDim3,blockDim3>, src/blas/dot.cuh:84-107.
 * The sum order is fixed by (gridDim3, blockDim3); the other template arguments do not change the result. */
size_t mpres_dot_workspace(const mpres_ctx_t *ctx, int n, int grid3, int block3);
int mpres_dot(mpres_ctx_t *ctx, int n, const mp_array_t *x, int incx, const mp_array_t *y, int incy, mp_array_t *r,
              int grid3, int block3, void *ws, size_t ws_bytes, void *stream);
/* mpres_dot without any allocation: no counterpart in the reference. For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same result as mpres_dot; the partial sums live in a fixed device array of
 * MPRES_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends and shared by every call on
 * that GPU, so calls that overlap in time on one GPU race. A reduction shape whose partials
 * (mpres_dot_workspace) exceed that size returns MPRES_ERR_WORKSPACE. */
#define MPRES_NOALLOC_SCRATCH_BYTES (4 << 20)
int mpres_dot_noalloc(mpres_ctx_t *ctx, int n, const mp_array_t *x, int incx, const mp_array_t *y, int incy,
                      mp_array_t *r, int grid3, int block3, void *stream);
/* y = alpha*x + y ; cuda::mp_axpy<gridDim1,blockDim1,gridDim2>, src/blas/axpy.cuh:45-74 (order-free) */
int mpres_axpy(mpres_ctx_t *ctx, int n, const mp_array_t *alpha, const mp_array_t *x, int incx, mp_array_t *y,
               int incy, void *stream);
/* y = alpha*op(A)*x + beta*y ; cuda::mp_gemv<gridDim1,blockDim1,gridDim2,blockDim3>, src/blas/gemv.cuh:150-268.
 * The sum order of each output is fixed by blockDim3. */
size_t mpres_gemv_workspace(const mpres_ctx_t *ctx, int trans, int m, int n, int block3);
int mpres_gemv(mpres_ctx_t *ctx, int trans, int m, int n, const mp_array_t *alpha, const mp_array_t *A, int lda,
               const mp_array_t *x, int incx, const mp_array_t *beta, mp_array_t *y, int incy, int block3, void *ws,
               size_t ws_bytes, void *stream);
/* C = alpha*A*B + beta*C ; cuda::mp_gemm<...>, src/blas/gemm.cuh:69-167 (no-trans only, as the reference).
 * Workspace: the m*n dot-product sums as mp_float_t (ws NULL: the context's pool).
 * A->digits and B->digits must be 16-byte aligned (device allocations are), else MPRES_ERR_ARG. */
size_t mpres_gemm_workspace(const mpres_ctx_t *ctx, int m, int n, int k);
int mpres_gemm(mpres_ctx_t *ctx, int transa, int transb, int m, int n, int k, const mp_array_t *alpha,
               const mp_array_t *A, int lda, const mp_array_t *B, int ldb, const mp_array_t *beta, mp_array_t *C,
               int ldc, void *ws, size_t ws_bytes, void *stream);

/* C = alpha*op(A)*op(B) + beta*C in the precision form on INT8 MFMA (no counterpart in the
 * reference; the reference's gemm is cuda::mp_gemm, src/blas/gemm.cuh:69-167). Every row of
 * op(A) and column of op(B) shares one power-of-two exponent; the aligned residues are split
 * into 7-bit limbs and multiplied exactly on INT8 MFMA with INT32 accumulation, so the product
 * is one exact sum per entry, rounded once to the working precision (the reference rounds every
 * product and partial sum). Rows or columns whose aligned magnitudes exceed the bit budget that
 * keeps |sum| < M/4 are truncated by the excess bits (tile alignment rounding). The epilogue
 * is (sum*alpha) + (C*beta) with mp_mul / mp_add. Results are not bit-identical to mpres_gemm.
 * Requires 0 <= k <= MPRES_MFMA_MAX_K; transposes allowed; A->digits and B->digits 16-byte
 * aligned, as for mpres_gemm. */
#define MPRES_MFMA_MAX_K 16384
size_t mpres_gemm_mfma_workspace(const mpres_ctx_t *ctx, int m, int n, int k);
int mpres_gemm_mfma(mpres_ctx_t *ctx, int transa, int transb, int m, int n, int k, const mp_array_t *alpha,
                    const mp_array_t *A, int lda, const mp_array_t *B, int ldb, const mp_array_t *beta, mp_array_t *C,
                    int ldc, void *ws, size_t ws_bytes, void *stream);

/* ---- the rest of the v1 BLAS. Each reference element kernel triple (esi + digits + round)
 * is one pass here, run in the reference's order with the reference's caller buffers, so the
 * results and every intermediate the reference leaves in a buffer are the reference's. The
 * reference's silent returns on invalid increments or leading dimensions are MPRES_ERR_ARG. */
/* x = alpha x: cuda::mp_scal, src/blas/scal.cuh:43-61 (returns if n <= 0 or incx <= 0) */
int mpres_scal(mpres_ctx_t *ctx, int n, const mp_array_t *alpha, mp_array_t *x, int incx, void *stream);
/* plane rotation: cuda::mp_rot, src/blas/rot.cuh:49-100. buffer1 = s x, buffer2 = s y, then
 * x = c x and y = c y over the first n elements with increment 1 (the reference calls mp_scal
 * with increment 1 whatever incx and incy are), then x = x + buffer2, y = y - buffer1 with
 * incx, incy. buffer1, buffer2: n numbers each. */
int mpres_rot(mpres_ctx_t *ctx, int n, mp_array_t *x, int incx, mp_array_t *y, int incy, const mp_array_t *c,
              const mp_array_t *s, mp_array_t *buffer1, mp_array_t *buffer2, void *stream);
/* w = alpha x + beta y: cuda::mp_waxpby, src/blas/waxpby.cuh:51-89 (buffer: n numbers) */
int mpres_waxpby(mpres_ctx_t *ctx, int n, const mp_array_t *alpha, const mp_array_t *x, int incx,
                 const mp_array_t *beta, const mp_array_t *y, int incy, mp_array_t *w, int incw, mp_array_t *buffer,
                 void *stream);
/* w = w - alpha v, r = u . w: cuda::mp_axpy_dot, src/blas/axpydot.cuh:54-86. buffer: n numbers,
 * holds alpha v on return (the reference's dot then overwrites it with u .* w); the dot is
 * mpres_dot with (grid3, block3) and its workspace. */
int mpres_axpy_dot(mpres_ctx_t *ctx, int n, const mp_array_t *alpha, mp_array_t *w, int incw, const mp_array_t *v,
                   int incv, const mp_array_t *u, int incu, mp_array_t *r, mp_array_t *buffer, int grid3, int block3,
                   void *ws, size_t ws_bytes, void *stream);
/* r = sum |x_i|: cuda::mp_asum<gridDim1, blockDim1>, src/blas/asum.cuh:41-70 (the reduction
 * kernels src/mpreduct.cuh:120-192; the sum order is fixed by (grid1, block1)) */
size_t mpres_asum_workspace(const mpres_ctx_t *ctx, int n, int grid1, int block1);
int mpres_asum(mpres_ctx_t *ctx, int n, const mp_array_t *x, int incx, mp_array_t *r, int grid1, int block1, void *ws,
               size_t ws_bytes, void *stream);
/* r = ||x||_1 (mpres_asum) or max |x_i| (MPRES_INF_NORM): cuda::mp_norm<gridDim1, blockDim1>,
 * src/blas/norm.cuh:43-79 (src/mpreduct.cuh:199-286). Workspace: mpres_asum_workspace. */
int mpres_norm(mpres_ctx_t *ctx, int norm, int n, const mp_array_t *x, int incx, mp_array_t *r, int grid1, int block1,
               void *ws, size_t ws_bytes, void *stream);
/* C = alpha A + beta B: cuda::mp_ge_add, src/blas/geadd.cuh:58-114 (buffer: m n numbers, ld m) */
int mpres_ge_add(mpres_ctx_t *ctx, int m, int n, const mp_array_t *alpha, const mp_array_t *A, int lda,
                 const mp_array_t *beta, const mp_array_t *B, int ldb, mp_array_t *C, int ldc, mp_array_t *buffer,
                 void *stream);
/* B = alpha A + beta B: cuda::mp_ge_acc, src/blas/geacc.cuh:57-113 (buffer: m n numbers) */
int mpres_ge_acc(mpres_ctx_t *ctx, int m, int n, const mp_array_t *alpha, const mp_array_t *A, int lda,
                 const mp_array_t *beta, mp_array_t *B, int ldb, mp_array_t *buffer, void *stream);
/* A = D A (MPRES_LEFT) or A D (MPRES_RIGHT), D diagonal stored as a vector:
 * cuda::mp_ge_diag_scale, src/blas/gediagscale.cuh:54-91 */
int mpres_ge_diag_scale(mpres_ctx_t *ctx, int side, int m, int n, const mp_array_t *D, int incd, mp_array_t *A,
                        int lda, void *stream);
/* A = DL A DR: cuda::mp_ge_lr_scale, src/blas/gelrscale.cuh:56-93 (left scaling first) */
int mpres_ge_lr_scale(mpres_ctx_t *ctx, int m, int n, const mp_array_t *DL, int incdl, const mp_array_t *DR,
                      int incdr, mp_array_t *A, int lda, void *stream);
/* r = ||A||_1 (largest column sum of |a|) or ||A||_inf (largest row sum): cuda::mp_ge_norm
 * <gridDim1, blockDim1>, src/blas/genorm.cuh:142-173. The column (row) sums land in buffer
 * (n or m numbers) as in the reference; block1 fixes their sum order, (grid1, block1) the
 * maximum's. */
size_t mpres_ge_norm_workspace(const mpres_ctx_t *ctx, int norm, int m, int n, int grid1, int block1);
int mpres_ge_norm(mpres_ctx_t *ctx, int norm, int m, int n, const mp_array_t *A, int lda, mp_array_t *r,
                  mp_array_t *buffer, int grid1, int block1, void *ws, size_t ws_bytes, void *stream);
/* A = A + alpha x y^T: cuda::mp_ger, src/blas/ger.cuh:155-214. buffer1 = alpha y (n numbers),
 * buffer2 = x buffer1^T (m n numbers, ld m, its len must be m n as the reference's), A += buffer2. */
int mpres_ger(mpres_ctx_t *ctx, int m, int n, const mp_array_t *alpha, const mp_array_t *x, int incx,
              const mp_array_t *y, int incy, mp_array_t *A, int lda, mp_array_t *buffer1, mp_array_t *buffer2,
              void *stream);

/* ---- BLAS, v2 interface over mp_float_t arrays (src/blas/v2) ---- */
/* cuda::mp_dot<blocks,threads>(n,x,y,r), src/blas/v2/dot_v2.cuh:121-129 */
size_t mpres_dot_v2_workspace(const mpres_ctx_t *ctx, int n, int blocks, int threads);
int mpres_dot_v2(mpres_ctx_t *ctx, int n, const void *x, const void *y, void *r, int blocks, int threads, void *ws,
                 size_t ws_bytes, void *stream);
/* cuda::mp_axpy(n,alpha,x,y,r), src/blas/v2/axpy_v2.cuh:39-52 */
int mpres_axpy_v2(mpres_ctx_t *ctx, int n, const void *alpha, const void *x, const void *y, void *r, void *stream);
/* r = (-alpha) x + y: cuda::mp_maxpy, src/blas/v2/maxpy_v2.cuh (order-free) */
int mpres_maxpy_v2(mpres_ctx_t *ctx, int n, const void *alpha, const void *x, const void *y, void *r, void *stream);
/* r = x - y: cuda::mp_diff, src/blas/v2/diff_v2.cuh */
int mpres_diff_v2(mpres_ctx_t *ctx, int n, const void *x, const void *y, void *r, void *stream);
/* r[i] = x[i] * d[i] (d: device doubles): cuda::mp_prod_d, src/blas/v2/prod_v2.cuh */
int mpres_prod_d_v2(mpres_ctx_t *ctx, int n, const void *x, const double *d, void *r, void *stream);
/* r[i] = alpha x[i]: cuda::mp_scal (v2), src/blas/v2/scal_v2.cuh:37-48 */
int mpres_scal_v2(mpres_ctx_t *ctx, int n, const void *alpha, const void *x, void *r, void *stream);
/* r[i] = x[i] y[i]: cuda::mp_prod, src/blas/v2/prod_v2.cuh:37-43 */
int mpres_prod_v2(mpres_ctx_t *ctx, int n, const void *x, const void *y, void *r, void *stream);
/* The v2 matrix kernels below stride row and column together, so the reference touches every
 * element only when its launch covers the matrix; these functions always cover it. */
/* A = A + (alpha x) y^T: cuda::mp_ger (v2), src/blas/v2/ger_v2.cuh:49-68 */
int mpres_ger_v2(mpres_ctx_t *ctx, int m, int n, const void *alpha, const void *x, int incx, const void *y, int incy,
                 void *A, int lda, void *stream);
/* A = A + (x x^T) alpha on the uplo triangle: cuda::mp_syr, src/blas/v2/syr_v2.cuh:58-76 */
int mpres_syr_v2(mpres_ctx_t *ctx, int uplo, int n, const void *alpha, const void *x, int incx, void *A, int lda,
                 void *stream);
/* A = A + ((x y^T) alpha + (y x^T) alpha) on the uplo triangle: cuda::mp_syr2,
 * src/blas/v2/syr2_v2.cuh:59-83 */
int mpres_syr2_v2(mpres_ctx_t *ctx, int uplo, int n, const void *alpha, const void *x, int incx, const void *y,
                  int incy, void *A, int lda, void *stream);
/* y = beta y + alpha (A x), A symmetric from its uplo triangle: cuda::mp_symv<threads>,
 * src/blas/v2/symv_v2.cuh:57-95 (each row summed in column order) */
int mpres_symv_v2(mpres_ctx_t *ctx, int uplo, int n, const void *alpha, const void *A, int lda, const void *x,
                  int incx, const void *beta, void *y, int incy, void *stream);
/* C = (op(A) op(A)^T) alpha + C beta on the uplo triangle: cuda::mp_syrk,
 * src/blas/v2/syrk_v2.cuh:61-85 (only MPRES_TRANS reads A transposed, as the reference) */
int mpres_syrk_v2(mpres_ctx_t *ctx, int uplo, int trans, int n, int k, const void *alpha, const void *A, int lda,
                  const void *beta, void *C, int ldc, void *stream);
/* *result (host) = ||x||_2: cuda::mp_norm2<blocks, threads>, src/blas/v2/norm2_v2.cuh:107-119
 * (sum of squares in the order of its two reduction kernels, then mp_dsqrt). Synchronizes
 * the stream. */
size_t mpres_norm2_v2_workspace(const mpres_ctx_t *ctx, int n, int blocks, int threads);
int mpres_norm2_v2(mpres_ctx_t *ctx, int n, const void *x, int blocks, int threads, double *result, void *ws,
                   size_t ws_bytes, void *stream);

/* ---- Matrix Market input (host functions, no GPU work): src/sparse/utils/mtx_reader.cuh ----
 * mpres_mtx_info: sizes, number of entry lines, longest row (maxnzr, mirrored entries of
 * symmetric files counted), diagonal entries (collect_mtx_stats :52-90); mpres_mtx_nnz: the
 * stored entries once mirrored. The builders fill caller-allocated host arrays exactly as the
 * reference's build_csr / build_ell / build_jad / build_dia / build_dense do, including the
 * within-row order left by its std::sort by row (which fixes the SpMV summation order). */
int mpres_mtx_info(const char *path, int *m, int *n, int *lines, int *maxnzr, int *nzmd, int *symmetric);
int mpres_mtx_nnz(int lines, int nzmd, int symmetric);
int mpres_mtx_ndiag(const char *path, int lines, int symmetric, int *ndiag);
int mpres_mtx_csr(const char *path, int m, int nnz, int lines, int symmetric, double *as, int *ja, int *irp);
int mpres_mtx_ell(const char *path, int m, int maxnzr, int lines, int symmetric, double *as, int *ja);
int mpres_mtx_jad(const char *path, int m, int maxnzr, int nnz, int lines, int symmetric, double *as, int *ja,
                  int *jcp, int *perm);
int mpres_mtx_dia(const char *path, int m, int lines, int symmetric, int ndiag, double *as, int *offset);
int mpres_mtx_dense(const char *path, int m, int n, int lines, int symmetric, double *a);
int mpres_read_vector(const char *path, int n, double *v);

/* ---- solvers ----
 * A x = b by conjugate gradients: mp_cg_csr<threads, blocks_reduce> (src/sparse/solver/
 * cg_csr.cuh:52-117), or with precond != NULL (device array of n doubles, the reciprocals of
 * A's diagonal) mp_pcg_csr with the diagonal preconditioner (pcg_csr.cuh:52-130). Each step
 * is the library's operation in the reference's order, so every iterate equals the
 * reference's. x: initial guess on entry (the reference assumes zero), solution on exit.
 * Stops when ||r_k|| <= tol ||r_0|| or after maxit iterations. resvec: host array of at
 * least maxit + 1 doubles, receives ||r_k|| / ||r_0|| for k = 0 .. *iters. The stopping test
 * reads the residual norm on the host, so the stream is synchronized once per iteration.
 * ws: mpres_cg_workspace() bytes, or NULL for the context's pool. */
size_t mpres_cg_workspace(const mpres_ctx_t *ctx, int n, int threads, int blocks_reduce);
/* a sparse matrix for mpres_cg: one of the formats above with its sizes */
#define MPRES_SPARSE_CSR 0
#define MPRES_SPARSE_ELL 1
#define MPRES_SPARSE_JAD 2
#define MPRES_SPARSE_DIA 3
typedef struct {
    int format;  /* MPRES_SPARSE_* */
    int maxnzr;  /* ELL, JAD */
    int ndiag;   /* DIA (the matrix is n x n) */
    csr_t csr;
    ell_t ell;
    jad_t jad;
    dia_t dia;
} mpres_sparse_t;
/* mp_cg_csr / mp_cg_ell / mp_cg_jad / mp_cg_dia and the mp_pcg_* with the diagonal
 * preconditioner (src/sparse/solver/{cg,pcg}_{csr,ell,jad,dia}.cuh): the solver above over A's
 * scalar SpMV kernel */
int mpres_cg(mpres_ctx_t *ctx, int n, const mpres_sparse_t *A, const void *b, double tol, int maxit, void *x,
             const double *precond, double *resvec, int *iters, int threads, int blocks_reduce, void *ws,
             size_t ws_bytes, void *stream);
int mpres_cg_csr(mpres_ctx_t *ctx, int n, const csr_t *A, const void *b, double tol, int maxit, void *x,
                 const double *precond, double *resvec, int *iters, int threads, int blocks_reduce, void *ws,
                 size_t ws_bytes, void *stream);

/* cuda::mp_gemm(transa,transb,...), src/blas/v2/gemm_v2.cuh:55-81.
 * Workspace: the m*n sums, as for mpres_gemm (mpres_gemm_workspace). */
int mpres_gemm_v2(mpres_ctx_t *ctx, int transa, int transb, int m, int n, int k, const void *alpha, const void *A,
                  int lda, const void *B, int ldb, const void *beta, void *C, int ldc, void *ws, size_t ws_bytes,
                  void *stream);

/* ---- sparse: double CSR matrix times mp vector (x, y: mp_float_t arrays) ----
 * Over 16 and 32 moduli every format runs in the split form (scalar chains, then residues)
 * and takes 4 * (49 + 2 m) bytes of the context's pool for its row order and REDO row list; over
 * CSR the long rows' residues run on the context's second stream, joined back before return. */
/* cuda::mp_spmv_csr<threads>, src/sparse/spmv/spmv_csr.cuh:48-64 */
int mpres_spmv_csr(mpres_ctx_t *ctx, int m, const csr_t *csr, const void *x, void *y, void *stream);
/* cuda::mp_spmv_csrv<threads,threadsPerRow>, src/sparse/spmv/spmv_csrv.cuh:48-87; threads_per_row in {1,2,4,8,16,32} */
int mpres_spmv_csrv(mpres_ctx_t *ctx, int m, const csr_t *csr, int threads_per_row, const void *x, void *y,
                    void *stream);
/* y = A x with A in ELLPACK form: mp_spmv_ell<threads>, src/sparse/spmv/spmv_ell.cuh:50-66.
 * ell->as, ell->ja: m x maxnzr column-major (slot c of row r at c*m + r); zero values are
 * skipped. Order: the slots of each row in turn. */
int mpres_spmv_ell(mpres_ctx_t *ctx, int m, int maxnzr, const ell_t *ell, const void *x, void *y, void *stream);
/* y = A x with A in jagged-diagonal form: mp_spmv_jad<threads> (src/sparse/spmv/spmv_jad.cuh:50-66)
 * and mp_spmv_jadv<threads, threadsPerRow> (spmv_jadv.cuh:50-87: lane-strided diagonals, then the
 * 16/8/4/2/1 tree); row r's result goes to y[perm[r]]. */
int mpres_spmv_jad(mpres_ctx_t *ctx, int m, int maxnzr, const jad_t *jad, const void *x, void *y, void *stream);
int mpres_spmv_jadv(mpres_ctx_t *ctx, int m, int maxnzr, const jad_t *jad, int threadsPerRow, const void *x, void *y,
                    void *stream);
/* y = A x with A in diagonal form: mp_spmv_dia<threads>, src/sparse/spmv/spmv_dia.cuh:50-66.
 * dia->as: m x ndiag column-major (diagonal i of row r at i*m + r), dia->offset: ndiag column
 * offsets; n: the number of columns. Terms with a column outside [0, n) or a zero value are
 * skipped. */
int mpres_spmv_dia(mpres_ctx_t *ctx, int m, int n, int ndiag, const dia_t *dia, const void *x, void *y, void *stream);

/* ---- batched form: many independent dot / gemv problems in one launch ----
 * No counterpart in the reference: a batch stands for a loop of the single calls
 * mpres_dot (cuda::mp_dot, src/blas/dot.cuh:84-107) and mpres_gemv (cuda::mp_gemv,
 * src/blas/gemv.cuh:150-268), and every problem's result is bit-identical to its
 * single call with the same order arguments. A plan holds a device table of work
 * items (whole dots and runs of gemv output rows, largest first) and a HIP graph of
 * one persistent kernel whose workgroups take items from a device counter;
 * mpres_batch_run launches that graph. The plan keeps the operands' device
 * pointers: the arrays must stay allocated while the plan is used, and their
 * contents may change between runs. No output (dot r, gemv y) may overlap any
 * operand of another problem or of its own problem: problems run in no fixed
 * order. */
#define MPRES_BATCH_DOT 0
#define MPRES_BATCH_GEMV 1
typedef struct {
    int op;               /* MPRES_BATCH_DOT or MPRES_BATCH_GEMV */
    int trans;            /* gemv: MPRES_NO_TRANS, MPRES_TRANS or MPRES_CONJ_TRANS */
    int m, n;             /* dot: length n (m unused); gemv: A is m x n (column-major) */
    int lda, incx, incy;  /* gemv: lda, incx, incy; dot: incx, incy */
    int grid3, block3;    /* dot: the reduction shape of mpres_dot; gemv: block3 of mpres_gemv */
    const mp_array_t *alpha, *A, *x, *beta;  /* gemv: alpha, A, x, beta; dot: x */
    mp_array_t *y;        /* dot: second vector (read); gemv: y (updated) */
    mp_array_t *r;        /* dot: result (one number); gemv: unused */
} mpres_batch_problem_t;
/* Geometry of the persistent kernel. It never changes results; 0 = built-in choice (the
 * gemv lanes of the context, 4 workgroups per CU, 128 rows: the autotuner's best on 4096
 * dot/gemv problems at 16 moduli, mpres-blas_amd/tuning/batch_gfx950.json). */
typedef struct {
    int lanes;            /* lanes per number (a value mpres_lanes_supported accepts) */
    int wg_per_cu;        /* persistent workgroups per compute unit */
    int gemv_rows;        /* output rows of a gemv per work item */
} mpres_batch_tuning_t;
typedef struct mpres_batch mpres_batch_t;
/* scratch bytes of a plan of these problems (0 for an empty batch) */
size_t mpres_batch_workspace(const mpres_ctx_t *ctx, const mpres_batch_problem_t *probs, int count,
                             const mpres_batch_tuning_t *tuning);
/* Checks every problem as its single call does (MPRES_ERR_ARG names the problem), builds the
 * item table and the graph. ws == NULL: the plan allocates its own scratch. tuning may be NULL. */
int mpres_batch_plan(mpres_ctx_t *ctx, const mpres_batch_problem_t *probs, int count,
                     const mpres_batch_tuning_t *tuning, void *ws, size_t ws_bytes, mpres_batch_t **plan);
int mpres_batch_run(mpres_batch_t *plan, void *stream);
/* multiply-accumulates of the batch (dot: n, gemv: m * n, summed) and the plan's item count */
int mpres_batch_info(const mpres_batch_t *plan, double *mp_fmas, int *items, mpres_batch_tuning_t *used);
int mpres_batch_destroy(mpres_batch_t *plan);

/* ---- timing of each operation's main kernel ----
 * With profiling on, the library brackets the main kernels of every dot (first
 * reduction pass), gemv (strided row sums), gemm (sums and epilogue), SpMV and axpy call with HIP
 * events on the call's stream. mpres_profile_read waits for the recorded
 * events and returns the accumulated kernel time and launch count of one
 * operation class (same op numbering as mpres_set_lanes; op 6: the residue
 * products of gemm over 16 and 32 moduli, whose op 3 time is the scalar chains;
 * op 7: mpres_gemm_mfma, all its kernels; op 8: mpres_batch_run, the whole graph; op 9:
 * gemv over 16 and 32 moduli, its residue kernel (op 2 time, on a second stream) and its
 * scalar-chain kernel together, from the fork to the join). */
int mpres_profile(mpres_ctx_t *ctx, int enable);
int mpres_profile_read(mpres_ctx_t *ctx, int op, double *ms, long *launches, int reset);

/* ---- tuning ----
 * Lanes per multiple-precision number used by the kernels of an operation
 * (0 = the built-in choice). op: 0 elementwise, 1 dot, 2 gemv, 3 gemm, 4 spmv, 5 axpy,
 * 7 precision-form gemm (mpres_gemm_mfma).
 * Only values compiled into the library are accepted (mpres_lanes_supported).
 * The built-in choice for the 8_2double, 16_4double and 32_8double sets is a table by call
 * size measured on MI355X (mpres-blas_amd/tuning/lanes_gfx950.json) for elementwise, dot,
 * gemv, gemm, spmv and axpy; mpres_set_lanes with a nonzero count replaces it by that fixed count,
 * with 0 restores it. mpres_get_lanes: the choice without a table. */
int mpres_lanes_supported(const mpres_ctx_t *ctx, int lanes);
int mpres_set_lanes(mpres_ctx_t *ctx, int op, int lanes);
int mpres_get_lanes(const mpres_ctx_t *ctx, int op);  /* lanes in use, or a negative error */
/* A table of lane counts by call size, standing in for the reference's compile-time launch
 * constants (src/kernel_config.cuh) chosen per size: a call of operation op whose size is at
 * most bounds[i] (the first such i; bounds increasing) uses lanes[i], a larger one
 * lanes[count - 1]; count = 0 removes the table (mpres_set_lanes' choice applies again).
 * Size: the call's multiply-accumulate count (dot and axpy n, gemv m n, gemm and the
 * precision form m n k), the row count for SpMV, the element count for mpres_elementwise.
 * No reference counterpart; results do not depend on it. mpres-blas_amd/lanes_tune.py
 * measures a table. mpres_get_lanes_for: the lanes a call of that size uses. */
#define MPRES_LANES_TABLE_MAX 8
int mpres_set_lanes_table(mpres_ctx_t *ctx, int op, int count, const long long *bounds, const int *lanes);
int mpres_get_lanes_for(const mpres_ctx_t *ctx, int op, long long size);

#ifdef __cplusplus
}
#endif
#endif
