"""Python binding of libmpres_hip.so (the C ABI in include/mpres_hip.h).

Thin ctypes layer used by the tests and the benchmark. Device memory comes
from hipMalloc (through the HIP runtime the library links) or from the
library's own mp_array_t allocator; every computation runs in the HIP library. Loading fails loudly
when the library has not been built: there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpres_hip.so")

MPRES_OK = 0
MPRES_ERR_ARG, MPRES_ERR_HIP, MPRES_ERR_WORKSPACE, MPRES_ERR_UNSUPPORTED = -1, -2, -3, -4
NO_TRANS, TRANS, CONJ_TRANS = 111, 112, 113
OP_ADD, OP_SUB, OP_MUL, OP_MUL_D, OP_VADD = 0, 1, 2, 3, 4
# operation classes of mpres_set_lanes
K_ELEM, K_DOT, K_GEMV, K_GEMM, K_SPMV, K_AXPY, K_MFMA = 0, 1, 2, 3, 4, 5, 7

ER_DTYPE = np.dtype([("frac", "<f8"), ("exp", "<i8")])


def mp_dtype(n):
    """mp_float_t with n moduli, src/types.cuh:69-74 (4n+40 bytes)."""
    return np.dtype([("digits", "<i4", (n,)), ("sign", "<i4"), ("exp", "<i4"), ("eval", ER_DTYPE, (2,))])


class MpArray(C.Structure):  # mp_array_t, src/types.cuh:85-92
    _fields_ = [("digits", C.c_void_p), ("sign", C.c_void_p), ("exp", C.c_void_p), ("eval", C.c_void_p),
                ("buf", C.c_void_p), ("len", C.c_void_p)]


class Jad(C.Structure):  # jad_t, src/types.cuh:118-123
    _fields_ = [("as_", C.c_void_p), ("ja", C.c_void_p), ("jcp", C.c_void_p), ("perm", C.c_void_p)]


class Ell(C.Structure):  # ell_t, src/types.cuh:128-131
    _fields_ = [("as_", C.c_void_p), ("ja", C.c_void_p)]


class Dia(C.Structure):  # dia_t, src/types.cuh:136-139
    _fields_ = [("as_", C.c_void_p), ("offset", C.c_void_p)]


class Csr(C.Structure):  # csr_t, src/types.cuh:109-113
    _fields_ = [("as_", C.c_void_p), ("ja", C.c_void_p), ("irp", C.c_void_p)]


class Sparse(C.Structure):  # mpres_sparse_t: a matrix of one format for mpres_cg
    _fields_ = [("format", C.c_int), ("maxnzr", C.c_int), ("ndiag", C.c_int), ("csr", Csr), ("ell", Ell),
                ("jad", Jad), ("dia", Dia)]


SPARSE_CSR, SPARSE_ELL, SPARSE_JAD, SPARSE_DIA = 0, 1, 2, 3
# mblas_enum.cuh values (src/blas/mblas_enum.cuh:31-44)
UPPER, LOWER = 121, 122
LEFT, RIGHT = 141, 142
ONE_NORM, INF_NORM = 171, 175

BATCH_DOT, BATCH_GEMV = 0, 1


class BatchProblem(C.Structure):  # mpres_batch_problem_t
    _fields_ = [("op", C.c_int), ("trans", C.c_int), ("m", C.c_int), ("n", C.c_int), ("lda", C.c_int),
                ("incx", C.c_int), ("incy", C.c_int), ("grid3", C.c_int), ("block3", C.c_int),
                ("alpha", C.c_void_p), ("A", C.c_void_p), ("x", C.c_void_p), ("beta", C.c_void_p),
                ("y", C.c_void_p), ("r", C.c_void_p)]


class BatchTuning(C.Structure):  # mpres_batch_tuning_t
    _fields_ = [("lanes", C.c_int), ("wg_per_cu", C.c_int), ("gemv_rows", C.c_int)]


def dot_problem(n, x, y, r, grid3, block3, incx=1, incy=1):
    """batch entry standing for mpres_dot(n, x, incx, y, incy, r, grid3, block3); x, y, r: MpArray"""
    return BatchProblem(op=BATCH_DOT, n=n, incx=incx, incy=incy, grid3=grid3, block3=block3,
                        x=C.addressof(x), y=C.addressof(y), r=C.addressof(r))


def gemv_problem(trans, m, n, alpha, A, lda, x, beta, y, block3, incx=1, incy=1):
    """batch entry standing for mpres_gemv(trans, m, n, alpha, A, lda, x, incx, beta, y, incy, block3)"""
    return BatchProblem(op=BATCH_GEMV, trans=trans, m=m, n=n, lda=lda, incx=incx, incy=incy, block3=block3,
                        alpha=C.addressof(alpha), A=C.addressof(A), x=C.addressof(x), beta=C.addressof(beta),
                        y=C.addressof(y))


class MpresError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MpresError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(LIB_PATH)
        vp, i32, sz = C.c_void_p, C.c_int, C.c_size_t
        L.mpres_last_error.restype = C.c_char_p
        L.mpres_moduli_set_name.restype = C.c_char_p
        L.mpres_moduli_set_id.argtypes = [C.c_char_p]
        L.mpres_ctx_create.argtypes = [C.POINTER(vp), i32, i32]
        L.mpres_ctx_destroy.argtypes = [vp]
        L.mpres_ctx_info.argtypes = [vp] + [C.POINTER(i32)] * 5
        L.mpres_mp_float_size.restype = sz
        L.mpres_mp_float_size.argtypes = [vp]
        L.mpres_ctx_table.argtypes = [vp, i32, vp, sz]
        L.mpres_moduli_set_table.argtypes = [i32, i32, vp, sz]
        L.mpres_array_init.argtypes = [vp, vp, sz]
        L.mpres_array_clear.argtypes = [vp, vp]
        L.mpres_array_host2device.argtypes = [vp, vp, vp, sz, vp]
        L.mpres_array_device2host.argtypes = [vp, vp, vp, sz, vp]
        L.mpres_array_set_d.argtypes = [vp, vp, vp, i32, vp]
        L.mpres_set_d.argtypes = [vp, vp, vp, i32, vp]
        L.mpres_aos_to_array.argtypes = [vp, vp, vp, i32, vp]
        L.mpres_array_to_aos.argtypes = [vp, vp, vp, i32, vp]
        L.mpres_elementwise.argtypes = [vp, i32, i32, vp, vp, vp, vp]
        L.mpres_cmp.argtypes = [vp, i32, i32, vp, vp, vp, vp]
        L.mpres_get_d.argtypes = [vp, i32, vp, vp, vp]
        L.mpres_div.argtypes = [vp, i32, vp, vp, vp, vp]
        L.mpres_sqrt.argtypes = [vp, i32, vp, vp, vp]
        L.mpres_dsqrt.argtypes = [vp, i32, vp, vp, vp]
        L.mpres_maxpy_v2.argtypes = [vp, i32, vp, vp, vp, vp, vp]
        L.mpres_diff_v2.argtypes = [vp, i32, vp, vp, vp, vp]
        L.mpres_prod_d_v2.argtypes = [vp, i32, vp, vp, vp, vp]
        L.mpres_norm2_v2_workspace.restype = sz
        L.mpres_norm2_v2_workspace.argtypes = [vp, i32, i32, i32]
        L.mpres_norm2_v2.argtypes = [vp, i32, vp, i32, i32, C.POINTER(C.c_double), vp, sz, vp]
        L.mpres_spmv_ell.argtypes = [vp, i32, i32, vp, vp, vp, vp]
        L.mpres_spmv_jad.argtypes = [vp, i32, i32, vp, vp, vp, vp]
        L.mpres_spmv_jadv.argtypes = [vp, i32, i32, vp, i32, vp, vp, vp]
        L.mpres_spmv_dia.argtypes = [vp, i32, i32, i32, vp, vp, vp, vp]
        L.mpres_cg.argtypes = [vp, i32, vp, vp, C.c_double, i32, vp, vp, C.POINTER(C.c_double), C.POINTER(i32), i32,
                               i32, vp, sz, vp]
        ip = C.POINTER(i32)
        L.mpres_mtx_info.argtypes = [C.c_char_p] + [ip] * 6
        L.mpres_mtx_nnz.argtypes = [i32, i32, i32]
        L.mpres_mtx_ndiag.argtypes = [C.c_char_p, i32, i32, ip]
        L.mpres_mtx_csr.argtypes = [C.c_char_p, i32, i32, i32, i32, vp, vp, vp]
        L.mpres_mtx_ell.argtypes = [C.c_char_p, i32, i32, i32, i32, vp, vp]
        L.mpres_mtx_jad.argtypes = [C.c_char_p, i32, i32, i32, i32, i32, vp, vp, vp, vp]
        L.mpres_mtx_dia.argtypes = [C.c_char_p, i32, i32, i32, i32, vp, vp]
        L.mpres_mtx_dense.argtypes = [C.c_char_p, i32, i32, i32, i32, vp]
        L.mpres_read_vector.argtypes = [C.c_char_p, i32, vp]
        L.mpres_cg_workspace.restype = sz
        L.mpres_cg_workspace.argtypes = [vp, i32, i32, i32]
        L.mpres_cg_csr.argtypes = [vp, i32, vp, vp, C.c_double, i32, vp, vp, C.POINTER(C.c_double),
                                   C.POINTER(i32), i32, i32, vp, sz, vp]
        L.mpres_array_get_d.argtypes = [vp, i32, vp, vp, vp]
        L.mpres_export_limbs.argtypes = [vp]
        L.mpres_get_integer.argtypes = [vp, i32, vp, vp, vp, vp, vp]
        L.mpres_rns_cmp.argtypes = [vp, i32, vp, vp, vp, vp]
        L.mpres_dot_workspace.restype = sz
        L.mpres_dot_workspace.argtypes = [vp, i32, i32, i32]
        L.mpres_dot.argtypes = [vp, i32, vp, i32, vp, i32, vp, i32, i32, vp, sz, vp]
        L.mpres_dot_noalloc.argtypes = [vp, i32, vp, i32, vp, i32, vp, i32, i32, vp]
        L.mpres_dot_v2_workspace.restype = sz
        L.mpres_dot_v2_workspace.argtypes = [vp, i32, i32, i32]
        L.mpres_dot_v2.argtypes = [vp, i32, vp, vp, vp, i32, i32, vp, sz, vp]
        L.mpres_axpy.argtypes = [vp, i32, vp, vp, i32, vp, i32, vp]
        L.mpres_axpy_v2.argtypes = [vp, i32, vp, vp, vp, vp, vp]
        L.mpres_gemv_workspace.restype = sz
        L.mpres_gemv_workspace.argtypes = [vp, i32, i32, i32, i32]
        L.mpres_gemv.argtypes = [vp, i32, i32, i32, vp, vp, i32, vp, i32, vp, vp, i32, i32, vp, sz, vp]
        L.mpres_gemm_workspace.restype = sz
        L.mpres_gemm_workspace.argtypes = [vp, i32, i32, i32]
        L.mpres_gemm.argtypes = [vp, i32, i32, i32, i32, i32, vp, vp, i32, vp, i32, vp, vp, i32, vp, sz, vp]
        L.mpres_gemm_mfma_workspace.restype = sz
        L.mpres_gemm_mfma_workspace.argtypes = [vp, i32, i32, i32]
        L.mpres_gemm_mfma.argtypes = [vp, i32, i32, i32, i32, i32, vp, vp, i32, vp, i32, vp, vp, i32, vp, sz, vp]
        L.mpres_gemm_v2.argtypes = [vp, i32, i32, i32, i32, i32, vp, vp, i32, vp, i32, vp, vp, i32, vp, sz, vp]
        L.mpres_spmv_csr.argtypes = [vp, i32, vp, vp, vp, vp]
        L.mpres_spmv_csrv.argtypes = [vp, i32, vp, i32, vp, vp, vp]
        L.mpres_batch_workspace.restype = sz
        L.mpres_batch_workspace.argtypes = [vp, vp, i32, vp]
        L.mpres_batch_plan.argtypes = [vp, vp, i32, vp, vp, sz, C.POINTER(vp)]
        L.mpres_batch_run.argtypes = [vp, vp]
        L.mpres_batch_info.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(i32), vp]
        L.mpres_batch_destroy.argtypes = [vp]
        # the rest of the v1 BLAS
        L.mpres_scal.argtypes = [vp, i32, vp, vp, i32, vp]
        L.mpres_rot.argtypes = [vp, i32, vp, i32, vp, i32, vp, vp, vp, vp, vp]
        L.mpres_waxpby.argtypes = [vp, i32, vp, vp, i32, vp, vp, i32, vp, i32, vp, vp]
        L.mpres_axpy_dot.argtypes = [vp, i32, vp, vp, i32, vp, i32, vp, i32, vp, vp, i32, i32, vp, sz, vp]
        L.mpres_asum_workspace.restype = sz
        L.mpres_asum_workspace.argtypes = [vp, i32, i32, i32]
        L.mpres_asum.argtypes = [vp, i32, vp, i32, vp, i32, i32, vp, sz, vp]
        L.mpres_norm.argtypes = [vp, i32, i32, vp, i32, vp, i32, i32, vp, sz, vp]
        L.mpres_ge_add.argtypes = [vp, i32, i32, vp, vp, i32, vp, vp, i32, vp, i32, vp, vp]
        L.mpres_ge_acc.argtypes = [vp, i32, i32, vp, vp, i32, vp, vp, i32, vp, vp]
        L.mpres_ge_diag_scale.argtypes = [vp, i32, i32, i32, vp, i32, vp, i32, vp]
        L.mpres_ge_lr_scale.argtypes = [vp, i32, i32, vp, i32, vp, i32, vp, i32, vp]
        L.mpres_ge_norm_workspace.restype = sz
        L.mpres_ge_norm_workspace.argtypes = [vp, i32, i32, i32, i32, i32]
        L.mpres_ge_norm.argtypes = [vp, i32, i32, i32, vp, i32, vp, vp, i32, i32, vp, sz, vp]
        L.mpres_ger.argtypes = [vp, i32, i32, vp, vp, i32, vp, i32, vp, i32, vp, vp, vp]
        # the rest of the v2 BLAS
        L.mpres_scal_v2.argtypes = [vp, i32, vp, vp, vp, vp]
        L.mpres_prod_v2.argtypes = [vp, i32, vp, vp, vp, vp]
        L.mpres_ger_v2.argtypes = [vp, i32, i32, vp, vp, i32, vp, i32, vp, i32, vp]
        L.mpres_syr_v2.argtypes = [vp, i32, i32, vp, vp, i32, vp, i32, vp]
        L.mpres_syr2_v2.argtypes = [vp, i32, i32, vp, vp, i32, vp, i32, vp, i32, vp]
        L.mpres_symv_v2.argtypes = [vp, i32, i32, vp, vp, i32, vp, i32, vp, vp, i32, vp]
        L.mpres_syrk_v2.argtypes = [vp, i32, i32, i32, i32, vp, vp, i32, vp, vp, i32, vp]
        L.mpres_profile.argtypes = [vp, i32]
        L.mpres_profile_read.argtypes = [vp, i32, C.POINTER(C.c_double), C.POINTER(C.c_long), i32]
        L.mpres_lanes_supported.argtypes = [vp, i32]
        L.mpres_set_lanes.argtypes = [vp, i32, i32]
        L.mpres_get_lanes.argtypes = [vp, i32]
        L.mpres_set_lanes_table.argtypes = [vp, i32, i32, C.POINTER(C.c_longlong), C.POINTER(C.c_int)]
        L.mpres_get_lanes_for.argtypes = [vp, i32, C.c_longlong]
        _lib = L
    return _lib


def check(rc):
    if rc != MPRES_OK:
        raise MpresError(f"mpres error {rc}: {lib().mpres_last_error().decode()}")
    return rc


def set_names():
    L = lib()
    return [L.mpres_moduli_set_name(i).decode() for i in range(L.mpres_num_moduli_sets())]


_hip = None


def hip():
    """The HIP runtime that libmpres_hip.so itself links (same process-wide instance).

    Device buffers and timing events are created through it directly; torch is
    not used here because it carries its own copy of the HIP runtime."""
    global _hip
    if _hip is None:
        lib()
        H = C.CDLL("libamdhip64.so.7")
        vp, sz = C.c_void_p, C.c_size_t
        H.hipMalloc.argtypes = [C.POINTER(vp), sz]
        H.hipFree.argtypes = [vp]
        H.hipMemcpy.argtypes = [vp, vp, sz, C.c_int]
        H.hipMemset.argtypes = [vp, C.c_int, sz]
        H.hipMemcpyAsync.argtypes = [vp, vp, sz, C.c_int, vp]
        H.hipEventCreate.argtypes = [C.POINTER(vp)]
        H.hipEventRecord.argtypes = [vp, vp]
        H.hipEventSynchronize.argtypes = [vp]
        H.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), vp, vp]
        H.hipEventDestroy.argtypes = [vp]
        H.hipStreamCreate.argtypes = [C.POINTER(vp)]
        H.hipStreamDestroy.argtypes = [vp]
        H.hipStreamSynchronize.argtypes = [vp]
        H.hipGetErrorString.restype = C.c_char_p
        _hip = H
    return _hip


def hipchk(rc):
    if rc != 0:
        raise MpresError(f"HIP error {rc}: {hip().hipGetErrorString(rc).decode()}")


def synchronize():
    hipchk(hip().hipDeviceSynchronize())


class DevBuf:
    """A device byte buffer (hipMalloc), optionally filled from a numpy array."""

    def __init__(self, nbytes=None, host=None):
        H = hip()
        if host is not None:
            host = np.ascontiguousarray(host)
            nbytes = host.nbytes
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        hipchk(H.hipMalloc(C.byref(p), max(self.nbytes, 1)))
        self.p = p
        if host is not None:
            hipchk(H.hipMemcpy(p, host.ctypes.data_as(C.c_void_p), self.nbytes, 1))
        else:
            hipchk(H.hipMemset(p, 0, max(self.nbytes, 1)))

    @property
    def ptr(self):
        return self.p

    def to_numpy(self, dtype, count=None):
        out = np.empty(self.nbytes, np.uint8)
        hipchk(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, self.nbytes, 2))
        a = out.view(dtype)
        return a if count is None else a[:count]

    def __del__(self):
        try:
            if self.p:
                hip().hipFree(self.p)
                self.p = None
        except Exception:
            pass


class Timer:
    """HIP events on the null stream (the stream the library's kernels use when no stream is given)."""

    def __init__(self):
        H = hip()
        self.a, self.b = C.c_void_p(), C.c_void_p()
        hipchk(H.hipEventCreate(C.byref(self.a)))
        hipchk(H.hipEventCreate(C.byref(self.b)))

    def start(self, stream=None):
        hipchk(hip().hipEventRecord(self.a, stream))

    def stop(self, stream=None):
        hipchk(hip().hipEventRecord(self.b, stream))

    def ms(self):
        hipchk(hip().hipEventSynchronize(self.b))
        t = C.c_float()
        hipchk(hip().hipEventElapsedTime(C.byref(t), self.a, self.b))
        return t.value


class Batch:
    """A batch plan (mpres_batch_plan) of BatchProblem entries; run() launches its graph.
    tuning: dict with keys of BatchTuning (lanes, wg_per_cu, gemv_rows), or None."""

    def __init__(self, ctx, problems, tuning=None):
        self.L = ctx.L
        arr = (BatchProblem * max(1, len(problems)))(*problems)
        t = C.byref(BatchTuning(**tuning)) if tuning else None
        h = C.c_void_p()
        check(self.L.mpres_batch_plan(ctx.h, arr, len(problems), t, None, 0, C.byref(h)))
        self.h = h

    def run(self, stream=None):
        check(self.L.mpres_batch_run(self.h, stream))

    def info(self):
        f, n, u = C.c_double(), C.c_int(), BatchTuning()
        check(self.L.mpres_batch_info(self.h, C.byref(f), C.byref(n), C.byref(u)))
        return {"mp_fmas": f.value, "items": n.value,
                "tuning": {"lanes": u.lanes, "wg_per_cu": u.wg_per_cu, "gemv_rows": u.gemv_rows}}

    def close(self):
        if self.h:
            self.L.mpres_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context:
    """One moduli set on one device (mpres_ctx_t)."""

    def __init__(self, set_name, device=0):
        L = lib()
        self.L = L
        sid = L.mpres_moduli_set_id(set_name.encode())
        if sid < 0:
            raise MpresError(f"unknown moduli set {set_name}")
        h = C.c_void_p()
        check(L.mpres_ctx_create(C.byref(h), sid, device))
        self.h = h
        vals = [C.c_int() for _ in range(5)]
        check(L.mpres_ctx_info(h, *[C.byref(v) for v in vals]))
        self.n, self.precision, self.mp_h, self.mp_j, self.log2m = [v.value for v in vals]
        self.size = L.mpres_mp_float_size(h)
        self.dtype = mp_dtype(self.n)
        self.set_name = set_name
        self._arrays = []

    def close(self):
        if self.h:
            for a in self._arrays:
                self.L.mpres_array_clear(self.h, C.byref(a))
            self._arrays = []
            self.L.mpres_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def profile(self, enable=True):
        check(self.L.mpres_profile(self.h, int(enable)))

    def profile_read(self, op, reset=True):
        ms, cnt = C.c_double(), C.c_long()
        check(self.L.mpres_profile_read(self.h, op, C.byref(ms), C.byref(cnt), int(reset)))
        return ms.value, cnt.value

    def set_lanes(self, op, lanes):
        check(self.L.mpres_set_lanes(self.h, op, lanes))

    def set_lanes_table(self, op, table):
        """table: [(bound, lanes), ...] with increasing bounds; [] removes the table"""
        n = len(table)
        b = (C.c_longlong * max(n, 1))(*[int(t[0]) for t in table])
        l_ = (C.c_int * max(n, 1))(*[int(t[1]) for t in table])
        check(self.L.mpres_set_lanes_table(self.h, op, n, b, l_))

    def lanes_for(self, op, size):
        return self.L.mpres_get_lanes_for(self.h, op, int(size))

    # ---- mp_array_t (SoA) ----
    def array(self, size):
        a = MpArray()
        check(self.L.mpres_array_init(self.h, C.byref(a), size))
        self._arrays.append(a)
        return a

    def array_from_host(self, recs):
        recs = np.ascontiguousarray(recs, dtype=self.dtype)
        a = self.array(len(recs))
        check(self.L.mpres_array_host2device(self.h, C.byref(a), recs.ctypes.data_as(C.c_void_p), len(recs), None))
        return a

    def copy_array(self, dst, src, count, stream=None):
        """Device copy of the first count numbers of one mp_array_t into another of the same length."""
        H = hip()
        n = self.n
        hipchk(H.hipMemcpyAsync(dst.digits, src.digits, 4 * n * count, 3, stream))
        hipchk(H.hipMemcpyAsync(dst.sign, src.sign, 4 * count, 3, stream))
        hipchk(H.hipMemcpyAsync(dst.exp, src.exp, 4 * count, 3, stream))
        hipchk(H.hipMemcpyAsync(dst.eval, src.eval, 32 * count, 3, stream))

    def array_to_host(self, a, count):
        out = np.zeros(count, self.dtype)
        check(self.L.mpres_array_device2host(self.h, out.ctypes.data_as(C.c_void_p), C.byref(a), count, None))
        return out

    # ---- mp_float_t arrays on the device (AoS) ----
    def aos(self, recs=None, count=None):
        if recs is not None:
            return DevBuf(host=np.ascontiguousarray(recs, dtype=self.dtype))
        return DevBuf(nbytes=self.size * count)

    def aos_to_host(self, buf, count):
        return buf.to_numpy(self.dtype, count).copy()

    # ---- conversions ----
    def set_d(self, x):
        """double -> mp on the GPU; returns the AoS device buffer."""
        x = np.ascontiguousarray(x, np.float64)
        dx = DevBuf(host=x)
        out = self.aos(count=len(x))
        check(self.L.mpres_set_d(self.h, out.ptr, dx.ptr, len(x), None))
        return out

    def array_set_d(self, x):
        x = np.ascontiguousarray(x, np.float64)
        dx = DevBuf(host=x)
        a = self.array(len(x))
        check(self.L.mpres_array_set_d(self.h, C.byref(a), dx.ptr, len(x), None))
        synchronize()
        return a

    def elementwise(self, op, x, y, count):
        r = self.aos(count=count)
        check(self.L.mpres_elementwise(self.h, op, count, x.ptr, y.ptr, r.ptr, None))
        return r

    def sync(self):
        synchronize()
