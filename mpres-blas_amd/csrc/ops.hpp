// Type-erased launcher table shared between the kernel translation units and
// the C ABI (api.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include "devconsts.hpp"

namespace mph {

struct ErMem {  // er_float_t in memory: {double frac; long exp}
    double frac;
    long long exp;
};

// mp_array_t / mp_collection_t (structure of arrays), src/types.cuh:85-104
struct SoA {
    int *d;
    int *s;
    int *e;
    ErMem *ev;
    const int *lenp;  // device pointer to the actual length (mp_array_t.len)
    int len;          // filled from *lenp inside the kernels: eval upper bounds start at ev + len
};

// a sparse matrix in one of the storage formats of the SpMV kernels (kernels.hpp)
enum { SP_CSR = 0, SP_ELL = 1, SP_JAD = 2, SP_DIA = 3 };
constexpr int SP_NB = 16;  // length classes of the split-form SpMV's row grouping (kernels.hpp)
struct SpMat {
    const double *as;
    const int *ja;      // CSR, ELL, JAD: column indices
    const int *ptr;     // CSR: irp; JAD: jcp
    const int *perm;    // JAD: row permutation
    const int *offset;  // DIA: diagonal offsets
    int width;          // ELL, JAD: maxnzr; DIA: ndiag
    int ncols;          // DIA: columns of the matrix
};

// element maps of the rest of the BLAS (kernels_blas.hpp): the v1 operation and where its
// scalar factor comes from, and the v2 element operations
enum { V1_MUL = 0, V1_ADD = 1, V1_SUB = 2 };
enum { SF_SCALAR = 0, SF_ROW = 1, SF_COL = 2 };
enum { V2_SCAL = 0, V2_GER = 1, V2_SYR = 2, V2_SYR2 = 3 };

// One unit of work of the batched form (mpres_batch_run), built on the host and
// read by the persistent kernel: a whole dot, or a run of output rows of a gemv.
enum BatchKind { BATCH_DOT = 0, BATCH_GEMV_N = 1, BATCH_GEMV_T = 2 };
struct BatchItem {
    int kind;
    int r0, r1;          // gemv: output rows [r0, r1) of this item
    int m, n;            // dot: n; gemv: A is m x n
    int lda, incx, incy;
    int g3, b3;          // dot: grid3 x block3; gemv: block3 (the summation order)
    long long ws;        // byte offset of the item's scratch in the batch workspace
    SoA alpha, A, x, beta, y, r;  // dot reads x, y and writes r; gemv reads alpha, A, x, beta and updates y
};

struct Ops {
    int N, L;
    hipError_t (*set_d)(const DevConsts *, const double *, int, SoA, char *, bool, hipStream_t);
    hipError_t (*elem)(const DevConsts *, int, int, const char *, const char *, const double *, char *, hipStream_t);
    // export: mp_get_d (mode 0) or the exact integers (mode 1) of count numbers, AoS (x) or SoA (so)
    hipError_t (*exp_out)(const DevConsts *, int, int, const char *, SoA, bool, double *, unsigned *, int *, int *,
                          hipStream_t);
    // comparisons (mode 0 mp_cmp, 1 mp_cmp_abs on AoS numbers, 2 rns_cmp on bare residues)
    hipError_t (*cmp)(const DevConsts *, int, int, const char *, const char *, const int *, const int *, int *,
                      hipStream_t);
    // the reduction launchers take a pass: 0 = main kernel, 1 = fixup kernel (see kernels.hpp);
    // gemv_fold runs passes 0..3 and gemm / gemm_v2 passes 0..4 (see launch.hpp)
    hipError_t (*fold)(const DevConsts *, int, int, long long, long long, SoA, int, SoA, int, const char *,
                       const char *, char *, hipStream_t);
    hipError_t (*tree)(const DevConsts *, char *, int, unsigned, int, char *, hipStream_t);
    hipError_t (*aos_to_soa)(const DevConsts *, const char *, SoA, int, hipStream_t);
    hipError_t (*soa_to_aos)(const DevConsts *, SoA, char *, int, hipStream_t);
    hipError_t (*axpy)(const DevConsts *, int, SoA, SoA, int, SoA, int, hipStream_t);
    hipError_t (*axpy_v2)(const DevConsts *, int, const char *, const char *, const char *, char *, bool, hipStream_t);
    // the reference's host MPFR operations: 0 mp_div (x / y), 1 mp_sqrt, 2 mp_dsqrt (into out)
    hipError_t (*hostop)(const DevConsts *, int, int, const char *, const char *, char *, double *, hipStream_t);
    hipError_t (*scal)(const DevConsts *, int, SoA, SoA, int, char *, bool, hipStream_t);
    hipError_t (*gemv_fold)(const DevConsts *, bool, int, int, int, SoA, int, const char *, int, long long, char *,
                            char *, hipStream_t);
    bool (*gemv_split)(bool trans, int b3);  // passes 1 and 2 of gemv_fold are independent
    size_t (*gemv_aux)(bool, int, int);
    hipError_t (*gemv_final)(const DevConsts *, int, SoA, int, const char *, hipStream_t);
    hipError_t (*gemm)(const DevConsts *, int, int, int, int, int, SoA, SoA, int, SoA, int, SoA, SoA, int, char *,
                       hipStream_t);
    hipError_t (*gemm_v2)(const DevConsts *, int, int, int, int, int, int, int, const char *, const char *, int,
                          const char *, int, const char *, char *, int, char *, hipStream_t);
    size_t (*gemm_ws)(int, int, int);
    hipError_t (*gemm_mfma)(const DevConsts *, int, int, int, int, int, int, int, SoA, SoA, int, SoA, int, SoA, SoA,
                            int, char *, hipStream_t);
    size_t (*gemm_mfma_ws)(int, int, int, int);
    // SpMV over a storage format (SP_CSR, SP_ELL, SP_JAD, SP_DIA of kernels.hpp): pass 0 main, 1 fixup,
    // 2 and 3 the short and long rows' residues of the split form over CSR (see launch.hpp)
    hipError_t (*spmv)(const DevConsts *, int, int, int, SpMat, const char *, char *, int, int *, hipStream_t);
    bool (*spmv_two)(int fmt);  // passes 2 and 3 have work (and may run at once on two streams)
    // the rest of the BLAS surface (kernels_blas.hpp)
    hipError_t (*v1map)(const DevConsts *, int, int, int, int, SoA, int, SoA, int, int, SoA, int, int, SoA, int, int,
                        hipStream_t);
    hipError_t (*fold_abs)(const DevConsts *, long long, long long, SoA, int, char *, hipStream_t);
    hipError_t (*maxabs)(const DevConsts *, int, long long, long long, SoA, int, const char *, char *, int, unsigned,
                         int, char *, hipStream_t);
    hipError_t (*matsum)(const DevConsts *, bool, int, int, SoA, int, int, char *, hipStream_t);
    hipError_t (*v2map)(const DevConsts *, int, int, int, int, const char *, const char *, int, const char *, int,
                        char *, int, char *, hipStream_t);
    hipError_t (*symv_v2)(const DevConsts *, int, int, const char *, const char *, int, const char *, int,
                          const char *, char *, int, hipStream_t);
    hipError_t (*syrk_v2)(const DevConsts *, int, int, int, int, const char *, const char *, int, const char *, char *,
                          int, hipStream_t);
    // batched form: reset the work queue, then the persistent kernel over `count` items on `grid` workgroups
    hipError_t (*batch)(const DevConsts *, const BatchItem *, int, int *, char *, int, hipStream_t);
};

// Compiled (moduli count, lanes per number) variants; each has its own
// translation unit csrc/k_n<N>_l<L>.hip defining ops_n<N>_l<L>().
#define MPH_VARIANTS(X) \
    X(8, 1) X(8, 2) X(8, 8) X(16, 1) X(16, 2) X(16, 4) X(16, 16) X(32, 1) X(32, 2) X(32, 4) X(32, 8) X(32, 32)

#define MPH_DECLARE(N, L) const Ops *ops_n##N##_l##L();
MPH_VARIANTS(MPH_DECLARE)
#undef MPH_DECLARE

// table of variant (N, L), or nullptr when it is not compiled
const Ops *find_variant(int N, int L);

}  // namespace mph
