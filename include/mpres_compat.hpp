// Header-only C++ layer that keeps the reference's entry-point names and
// template signatures (namespace cuda, src/blas/*.cuh, src/blas/v2/*.cuh,
// src/sparse/spmv/*.cuh) and forwards each call to the C ABI of
// mpres_hip.h. A program written against the reference's BLAS calls
// compiles against this header with its call sites unchanged, except that
// kernels the reference exposes as __global__ (v2 mp_axpy and mp_gemm,
// mp_spmv_csr(v)) become host functions: `f<T><<<g, b>>>(args)` becomes
// `f<T>(args)`.
//
// The moduli set is fixed at compile time by MPRES_COMPAT_MODULI_SET (the
// reference selects it by overwriting src/params.h); MPRES_COMPAT_N must
// match its number of moduli. Template arguments that fix a summation order
// are forwarded (dot gridDim3 x blockDim3, gemv blockDim3, CSR-vector
// threadsPerRow); the others are launch shapes of the reference's kernels
// and do not change results, so they are ignored. Errors throw
// std::runtime_error with the library's message (the reference prints and
// continues).
#ifndef MPRES_COMPAT_HPP
#define MPRES_COMPAT_HPP

#include <stdexcept>
#include <string>

#include "mpres_hip.h"

#ifndef MPRES_COMPAT_MODULI_SET
#define MPRES_COMPAT_MODULI_SET "8_2double"
#endif
#ifndef MPRES_COMPAT_N
#define MPRES_COMPAT_N 8
#endif

// mp_float_t with the compile-time number of moduli, src/types.cuh:69-76
typedef struct {
    int digits[MPRES_COMPAT_N];
    int sign;
    int exp;
    er_float_t eval[2];
} mp_float_t;
typedef mp_float_t *mp_float_ptr;

// src/blas/mblas_enum.cuh:25-29
enum mblas_trans_type { mblas_no_trans = MPRES_NO_TRANS, mblas_trans = MPRES_TRANS, mblas_conj_trans = MPRES_CONJ_TRANS };

namespace mpres_compat {

inline mpres_ctx_t *&context() {
    static mpres_ctx_t *ctx = nullptr;
    return ctx;
}
inline void check(int rc) {
    if (rc != MPRES_OK) throw std::runtime_error(std::string("mpres: ") + mpres_last_error());
}
// rns_const_init() + mp_const_init() (src/rns.cuh:324, src/arith/arith_utils.cuh:44)
inline void init(int device = 0) {
    if (context()) return;
    const int id = mpres_moduli_set_id(MPRES_COMPAT_MODULI_SET);
    if (id < 0) throw std::runtime_error("mpres: unknown moduli set " MPRES_COMPAT_MODULI_SET);
    check(mpres_ctx_create(&context(), id, device));
    int n = 0;
    check(mpres_ctx_info(context(), &n, nullptr, nullptr, nullptr, nullptr));
    if (n != MPRES_COMPAT_N) throw std::runtime_error("mpres: MPRES_COMPAT_N does not match the moduli set");
}
inline mpres_ctx_t *ctx() {
    init();
    return context();
}

}  // namespace mpres_compat

inline void rns_const_init() { mpres_compat::init(); }
inline void mp_const_init() { mpres_compat::init(); }

namespace cuda {

// src/mparray.cuh:35-165
inline void mp_array_init(mp_array_t &dev, size_t size) { mpres_compat::check(mpres_array_init(mpres_compat::ctx(), &dev, size)); }
inline void mp_array_clear(mp_array_t &dev) { mpres_compat::check(mpres_array_clear(mpres_compat::ctx(), &dev)); }
inline void mp_array_host2device(mp_array_t &dev, mp_float_ptr host, size_t size) {
    mpres_compat::check(mpres_array_host2device(mpres_compat::ctx(), &dev, host, size, nullptr));
}
inline void mp_array_device2host(mp_float_ptr host, mp_array_t &dev, size_t size) {
    mpres_compat::check(mpres_array_device2host(mpres_compat::ctx(), host, &dev, size, nullptr));
}

// src/blas/dot.cuh:84-107; buffer: the reference's product buffer (not needed here)
template <int gridDim1, int blockDim1, int gridDim2, int gridDim3, int blockDim3>
void mp_dot(const int n, mp_array_t &x, const int incx, mp_array_t &y, const int incy, mp_array_t &r,
            mp_array_t &buffer) {
    (void)buffer;
    mpres_compat::check(mpres_dot(mpres_compat::ctx(), n, &x, incx, &y, incy, &r, gridDim3, blockDim3, nullptr, 0,
                                  nullptr));
}

// src/blas/axpy.cuh:45-74
template <int gridDim1, int blockDim1, int gridDim2>
void mp_axpy(const int n, mp_array_t &alpha, mp_array_t &x, const int incx, mp_array_t &y, const int incy,
             mp_array_t &buffer) {
    (void)buffer;
    mpres_compat::check(mpres_axpy(mpres_compat::ctx(), n, &alpha, &x, incx, &y, incy, nullptr));
}

// src/blas/gemv.cuh:150-268
template <int gridDim1, int blockDim1, int gridDim2, int blockDim3>
void mp_gemv(enum mblas_trans_type trans, const int m, const int n, mp_array_t &alpha, mp_array_t &A, const int lda,
             mp_array_t &x, const int incx, mp_array_t &beta, mp_array_t &y, const int incy, mp_array_t &buffer1,
             mp_array_t &buffer2) {
    (void)buffer1;
    (void)buffer2;
    mpres_compat::check(mpres_gemv(mpres_compat::ctx(), trans, m, n, &alpha, &A, lda, &x, incx, &beta, &y, incy,
                                   blockDim3, nullptr, 0, nullptr));
}

// src/blas/gemm.cuh:69-167
template <int blockDim1x, int blockDim1y, int gridDim2x, int gridDim2y, int blockDim3>
void mp_gemm(enum mblas_trans_type transa, enum mblas_trans_type transb, const int m, const int n, const int k,
             mp_array_t &alpha, mp_array_t &A, const int lda, mp_array_t &B, const int ldb, mp_array_t &beta,
             mp_array_t &C, const int ldc, mp_array_t &buffer) {
    (void)buffer;
    mpres_compat::check(mpres_gemm(mpres_compat::ctx(), transa, transb, m, n, k, &alpha, &A, lda, &B, ldb, &beta, &C,
                                   ldc, nullptr, 0, nullptr));
}

// v2: src/blas/v2/dot_v2.cuh:121-129
template <int blocks, int threads>
void mp_dot(const int n, mp_float_ptr x, mp_float_ptr y, mp_float_ptr r) {
    mpres_compat::check(mpres_dot_v2(mpres_compat::ctx(), n, x, y, r, blocks, threads, nullptr, 0, nullptr));
}
// v2: src/blas/v2/axpy_v2.cuh:39-52 (a __global__ kernel in the reference)
inline void mp_axpy(const int n, mp_float_ptr alpha, mp_float_ptr x, mp_float_ptr y, mp_float_ptr r) {
    mpres_compat::check(mpres_axpy_v2(mpres_compat::ctx(), n, alpha, x, y, r, nullptr));
}
// v2: src/blas/v2/gemm_v2.cuh:55-81 (a __global__ kernel in the reference)
inline void mp_gemm(enum mblas_trans_type transa, enum mblas_trans_type transb, const int m, const int n, const int k,
                    mp_float_ptr alpha, mp_float_ptr A, const int lda, mp_float_ptr B, const int ldb, mp_float_ptr beta,
                    mp_float_ptr C, const int ldc) {
    mpres_compat::check(mpres_gemm_v2(mpres_compat::ctx(), transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C,
                                      ldc, nullptr, 0, nullptr));
}

// src/sparse/spmv/spmv_csr.cuh:48-64 (a __global__ kernel in the reference)
template <int threads>
void mp_spmv_csr(const int m, const csr_t csr, mp_float_ptr x, mp_float_ptr y) {
    mpres_compat::check(mpres_spmv_csr(mpres_compat::ctx(), m, &csr, x, y, nullptr));
}
// src/sparse/spmv/spmv_csrv.cuh:48-87 (a __global__ kernel in the reference)
template <int threads, int threadsPerRow>
void mp_spmv_csrv(const int m, const csr_t csr, mp_float_ptr x, mp_float_ptr y) {
    mpres_compat::check(mpres_spmv_csrv(mpres_compat::ctx(), m, &csr, threadsPerRow, x, y, nullptr));
}

}  // namespace cuda

#endif
