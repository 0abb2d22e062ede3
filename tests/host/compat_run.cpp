// Runs every reference-named entry point of include/mpres_compat.hpp on seeded inputs
// and compares each result, bit for bit, with the direct C-ABI call that the name must
// forward to (argument order, template values that fix a summation order, the v1/v2
// layouts). Built and run by tests/test_gpu_compat.py on a GPU; prints one line per
// check and exits non-zero on the first mismatch.
#define MPRES_COMPAT_MODULI_SET "16_4double"
#define MPRES_COMPAT_N 16
#include "mpres_compat.hpp"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

namespace {

int checks = 0;

void hipok(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(2);
    }
}

std::vector<double> uexp(unsigned seed, size_t n, int lo, int hi) {
    std::mt19937_64 g(seed);
    std::vector<double> v(n);
    for (auto &x : v) {
        const double u = 2.0 * (double)(g() >> 11) * 0x1p-53 - 1.0;
        x = std::ldexp(u, lo + (int)(g() % (unsigned)(hi - lo + 1)));
    }
    return v;
}

double *dev_doubles(const std::vector<double> &v) {
    double *p = nullptr;
    hipok(hipMalloc(&p, v.size() * sizeof(double)), "hipMalloc");
    hipok(hipMemcpy(p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy");
    return p;
}

// v1 array from doubles (mp_set_d on the GPU)
void soa_of(mp_array_t &a, const std::vector<double> &v) {
    cuda::mp_array_init(a, v.size());
    double *d = dev_doubles(v);
    mpres_compat::check(mpres_array_set_d(mpres_compat::ctx(), &a, d, (int)v.size(), nullptr));
    hipok(hipDeviceSynchronize(), "sync");
    hipok(hipFree(d), "hipFree");
}
// v2 (mp_float_t) device array from doubles
mp_float_ptr aos_of(const std::vector<double> &v) {
    mp_float_ptr p = nullptr;
    hipok(hipMalloc(&p, v.size() * sizeof(mp_float_t)), "hipMalloc");
    double *d = dev_doubles(v);
    mpres_compat::check(mpres_set_d(mpres_compat::ctx(), p, d, (int)v.size(), nullptr));
    hipok(hipDeviceSynchronize(), "sync");
    hipok(hipFree(d), "hipFree");
    return p;
}
mp_float_ptr aos_copy(mp_float_ptr src, size_t n) {
    mp_float_ptr p = nullptr;
    hipok(hipMalloc(&p, n * sizeof(mp_float_t)), "hipMalloc");
    hipok(hipMemcpy(p, src, n * sizeof(mp_float_t), hipMemcpyDeviceToDevice), "hipMemcpy");
    return p;
}
std::vector<mp_float_t> host_soa(mp_array_t &a, size_t n) {
    std::vector<mp_float_t> h(n);
    cuda::mp_array_device2host(h.data(), a, n);
    return h;
}
std::vector<mp_float_t> host_aos(mp_float_ptr p, size_t n) {
    std::vector<mp_float_t> h(n);
    hipok(hipMemcpy(h.data(), p, n * sizeof(mp_float_t), hipMemcpyDeviceToHost), "hipMemcpy");
    return h;
}
void same(const std::vector<mp_float_t> &a, const std::vector<mp_float_t> &b, const char *what) {
    if (a.size() != b.size() || std::memcmp(a.data(), b.data(), a.size() * sizeof(mp_float_t)) != 0) {
        std::fprintf(stderr, "MISMATCH %s\n", what);
        std::exit(1);
    }
    checks++;
    std::printf("ok %s (%zu records)\n", what, a.size());
    std::fflush(stdout);
}

}  // namespace

int main() {
    rns_const_init();
    mp_const_init();
    mpres_ctx_t *ctx = mpres_compat::ctx();
    mp_array_t buf{};  // the reference's scratch buffers; the compat layer needs none

    // ---- v1 dot: cuda::mp_dot<gridDim1, blockDim1, gridDim2, gridDim3, blockDim3> ----
    {
        const int n = 5000;
        mp_array_t x, y, r1, r2;
        soa_of(x, uexp(1, n, -20, 20));
        soa_of(y, uexp(2, n, -20, 20));
        cuda::mp_array_init(r1, 1);
        cuda::mp_array_init(r2, 1);
        cuda::mp_dot<64, 64, 32, 32, 64>(n, x, 1, y, 1, r1, buf);
        mpres_compat::check(mpres_dot(ctx, n, &x, 1, &y, 1, &r2, 32, 64, nullptr, 0, nullptr));
        same(host_soa(r1, 1), host_soa(r2, 1), "mp_dot<64,64,32,32,64>");
        // v1 axpy
        mp_array_t alpha, y1, y2;
        soa_of(alpha, uexp(3, 1, -20, 20));
        soa_of(y1, uexp(2, n, -20, 20));
        soa_of(y2, uexp(2, n, -20, 20));
        cuda::mp_axpy<64, 64, 32>(n, alpha, x, 1, y1, 1, buf);
        mpres_compat::check(mpres_axpy(ctx, n, &alpha, &x, 1, &y2, 1, nullptr));
        same(host_soa(y1, n), host_soa(y2, n), "mp_axpy<64,64,32>");
        for (mp_array_t *a : {&x, &y, &r1, &r2, &alpha, &y1, &y2}) cuda::mp_array_clear(*a);
    }
    // ---- v1 gemv, both transposes: cuda::mp_gemv<gridDim1, blockDim1, gridDim2, blockDim3> ----
    {
        const int m = 300, n = 200;
        for (int tr = 0; tr < 2; tr++) {
            const int lx = tr ? m : n, ly = tr ? n : m;
            mp_array_t A, x, alpha, beta, y1, y2;
            soa_of(A, uexp(11, (size_t)m * n, -30, 30));
            soa_of(x, uexp(12, lx, -30, 30));
            soa_of(alpha, uexp(13, 1, -30, 30));
            soa_of(beta, uexp(14, 1, -30, 30));
            soa_of(y1, uexp(15, ly, -30, 30));
            soa_of(y2, uexp(15, ly, -30, 30));
            cuda::mp_gemv<64, 64, 32, 16>(tr ? mblas_trans : mblas_no_trans, m, n, alpha, A, m, x, 1, beta, y1, 1, buf,
                                          buf);
            mpres_compat::check(mpres_gemv(ctx, tr ? MPRES_TRANS : MPRES_NO_TRANS, m, n, &alpha, &A, m, &x, 1, &beta,
                                           &y2, 1, 16, nullptr, 0, nullptr));
            same(host_soa(y1, ly), host_soa(y2, ly), tr ? "mp_gemv<...,16> trans" : "mp_gemv<...,16> no-trans");
            for (mp_array_t *a : {&A, &x, &alpha, &beta, &y1, &y2}) cuda::mp_array_clear(*a);
        }
    }
    // ---- v1 gemm: cuda::mp_gemm<blockDim1x, blockDim1y, gridDim2x, gridDim2y, blockDim3> ----
    {
        const int m = 70, n = 50, k = 40;
        mp_array_t A, B, alpha, beta, C1, C2;
        soa_of(A, uexp(21, (size_t)m * k, -30, 30));
        soa_of(B, uexp(22, (size_t)k * n, -30, 30));
        soa_of(alpha, uexp(23, 1, -30, 30));
        soa_of(beta, uexp(24, 1, -30, 30));
        soa_of(C1, uexp(25, (size_t)m * n, -30, 30));
        soa_of(C2, uexp(25, (size_t)m * n, -30, 30));
        cuda::mp_gemm<16, 16, 8, 8, 32>(mblas_no_trans, mblas_no_trans, m, n, k, alpha, A, m, B, k, beta, C1, m, buf);
        mpres_compat::check(mpres_gemm(ctx, MPRES_NO_TRANS, MPRES_NO_TRANS, m, n, k, &alpha, &A, m, &B, k, &beta, &C2,
                                       m, nullptr, 0, nullptr));
        same(host_soa(C1, (size_t)m * n), host_soa(C2, (size_t)m * n), "mp_gemm<16,16,8,8,32>");
        for (mp_array_t *a : {&A, &B, &alpha, &beta, &C1, &C2}) cuda::mp_array_clear(*a);
    }
    // ---- v2 (mp_float_t arrays): dot<blocks, threads>, axpy, gemm with transposes ----
    {
        const int n = 3000;
        mp_float_ptr x = aos_of(uexp(31, n, -20, 20)), y = aos_of(uexp(32, n, -20, 20));
        mp_float_ptr r1 = aos_of({0.0}), r2 = aos_of({0.0});
        cuda::mp_dot<32, 64>(n, x, y, r1);
        mpres_compat::check(mpres_dot_v2(ctx, n, x, y, r2, 32, 64, nullptr, 0, nullptr));
        same(host_aos(r1, 1), host_aos(r2, 1), "v2 mp_dot<32,64>");
        mp_float_ptr alpha = aos_of(uexp(33, 1, -20, 20));
        mp_float_ptr o1 = aos_of(std::vector<double>(n, 0.0)), o2 = aos_of(std::vector<double>(n, 0.0));
        cuda::mp_axpy(n, alpha, x, y, o1);
        mpres_compat::check(mpres_axpy_v2(ctx, n, alpha, x, y, o2, nullptr));
        same(host_aos(o1, n), host_aos(o2, n), "v2 mp_axpy");
        const int m = 40, nn = 30, k = 20;
        mp_float_ptr A = aos_of(uexp(34, (size_t)m * k, -30, 30)), B = aos_of(uexp(35, (size_t)k * nn, -30, 30));
        mp_float_ptr beta = aos_of(uexp(36, 1, -30, 30));
        mp_float_ptr C0 = aos_of(uexp(37, (size_t)m * nn, -30, 30));
        mp_float_ptr C1 = aos_copy(C0, (size_t)m * nn), C2 = aos_copy(C0, (size_t)m * nn);
        cuda::mp_gemm(mblas_trans, mblas_no_trans, m, nn, k, alpha, A, k, B, k, beta, C1, m);
        mpres_compat::check(mpres_gemm_v2(ctx, MPRES_TRANS, MPRES_NO_TRANS, m, nn, k, alpha, A, k, B, k, beta, C2, m,
                                          nullptr, 0, nullptr));
        same(host_aos(C1, (size_t)m * nn), host_aos(C2, (size_t)m * nn), "v2 mp_gemm (trans, no-trans)");
        for (mp_float_ptr p : {x, y, r1, r2, alpha, o1, o2, A, B, beta, C0, C1, C2}) hipok(hipFree(p), "hipFree");
    }
    // ---- SpMV: mp_spmv_csr<threads>, mp_spmv_csrv<threads, threadsPerRow> ----
    {
        const int m = 500;
        std::vector<int> irp(m + 1), ja;
        std::vector<double> as;
        std::mt19937_64 g(41);
        for (int i = 0; i < m; i++) {
            irp[i] = (int)ja.size();
            const int len = 1 + (int)(g() % 40);
            for (int q = 0; q < len; q++) {
                ja.push_back((int)(g() % m));
                as.push_back(2.0 * (double)(g() >> 11) * 0x1p-53 - 1.0);
            }
        }
        irp[m] = (int)ja.size();
        csr_t csr;
        hipok(hipMalloc(&csr.irp, irp.size() * sizeof(int)), "hipMalloc");
        hipok(hipMalloc(&csr.ja, ja.size() * sizeof(int)), "hipMalloc");
        csr.as = dev_doubles(as);
        hipok(hipMemcpy(csr.irp, irp.data(), irp.size() * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
        hipok(hipMemcpy(csr.ja, ja.data(), ja.size() * sizeof(int), hipMemcpyHostToDevice), "hipMemcpy");
        mp_float_ptr x = aos_of(uexp(42, m, -10, 10));
        mp_float_ptr y1 = aos_of(std::vector<double>(m, 0.0)), y2 = aos_of(std::vector<double>(m, 0.0));
        cuda::mp_spmv_csr<32>(m, csr, x, y1);
        mpres_compat::check(mpres_spmv_csr(ctx, m, &csr, x, y2, nullptr));
        same(host_aos(y1, m), host_aos(y2, m), "mp_spmv_csr<32>");
        cuda::mp_spmv_csrv<64, 8>(m, csr, x, y1);
        mpres_compat::check(mpres_spmv_csrv(ctx, m, &csr, 8, x, y2, nullptr));
        same(host_aos(y1, m), host_aos(y2, m), "mp_spmv_csrv<64,8>");
        for (void *p : {(void *)x, (void *)y1, (void *)y2, (void *)csr.irp, (void *)csr.ja, (void *)csr.as})
            hipok(hipFree(p), "hipFree");
    }
    std::printf("compat run ok: %d checks\n", checks);
    return 0;
}
