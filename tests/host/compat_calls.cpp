// Compile-and-link check of include/mpres_compat.hpp: every reference-named
// entry point, called the way the reference's own harnesses call it
// (tests/blas/performance/*.cu), must compile against the header and link
// against libmpres_hip.so. Built by tests/test_capi_cpu.py; not run.
#define MPRES_COMPAT_MODULI_SET "32_8double"
#define MPRES_COMPAT_N 32
#include "mpres_compat.hpp"
int main() {
    rns_const_init();
    mp_const_init();
    mp_array_t x, y, r, buf;
    cuda::mp_array_init(x, 10);
    cuda::mp_dot<64, 64, 32, 32, 64>(10, x, 1, y, 1, r, buf);
    cuda::mp_gemv<64, 64, 32, 32>(mblas_no_trans, 4, 4, x, x, 4, x, 1, x, y, 1, buf, buf);
    cuda::mp_gemm<16, 16, 8, 8, 32>(mblas_no_trans, mblas_no_trans, 2, 2, 2, x, x, 2, x, 2, x, y, 2, buf);
    cuda::mp_axpy<64, 64, 32>(10, x, x, 1, y, 1, buf);
    mp_float_ptr p = nullptr;
    cuda::mp_dot<64, 64>(10, p, p, p);
    cuda::mp_axpy(10, p, p, p, p);
    cuda::mp_gemm(mblas_trans, mblas_no_trans, 2, 2, 2, p, p, 2, p, 2, p, p, 2);
    csr_t csr{};
    cuda::mp_spmv_csr<32>(4, csr, p, p);
    cuda::mp_spmv_csrv<32, 4>(4, csr, p, p);
    static_assert(sizeof(mp_float_t) == 4 * 32 + 40, "layout");
    return 0;
}
