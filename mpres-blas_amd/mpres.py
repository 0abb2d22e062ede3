"""Python binding of libmpres_hip.so (the C ABI in include/mpres_hip.h).

Thin ctypes layer used by the tests and the benchmark. Device memory comes
from torch (plain byte tensors) or from the library's own mp_array_t
allocator; every computation runs in the HIP library. Loading fails loudly
when the library has not been built: there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpres_hip.so")

MPRES_OK = 0
NO_TRANS, TRANS, CONJ_TRANS = 111, 112, 113
OP_ADD, OP_SUB, OP_MUL, OP_MUL_D, OP_VADD = 0, 1, 2, 3, 4
# operation classes of mpres_set_lanes
K_ELEM, K_DOT, K_GEMV, K_GEMM, K_SPMV, K_AXPY = 0, 1, 2, 3, 4, 5

ER_DTYPE = np.dtype([("frac", "<f8"), ("exp", "<i8")])


def mp_dtype(n):
    """mp_float_t with n moduli, src/types.cuh:69-74 (4n+40 bytes)."""
    return np.dtype([("digits", "<i4", (n,)), ("sign", "<i4"), ("exp", "<i4"), ("eval", ER_DTYPE, (2,))])


class MpArray(C.Structure):  # mp_array_t, src/types.cuh:85-92
    _fields_ = [("digits", C.c_void_p), ("sign", C.c_void_p), ("exp", C.c_void_p), ("eval", C.c_void_p),
                ("buf", C.c_void_p), ("len", C.c_void_p)]


class Csr(C.Structure):  # csr_t, src/types.cuh:109-113
    _fields_ = [("as_", C.c_void_p), ("ja", C.c_void_p), ("irp", C.c_void_p)]


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
        L.mpres_dot_workspace.restype = sz
        L.mpres_dot_workspace.argtypes = [vp, i32, i32, i32]
        L.mpres_dot.argtypes = [vp, i32, vp, i32, vp, i32, vp, i32, i32, vp, sz, vp]
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
        L.mpres_gemm_v2.argtypes = [vp, i32, i32, i32, i32, i32, vp, vp, i32, vp, i32, vp, vp, i32, vp]
        L.mpres_spmv_csr.argtypes = [vp, i32, vp, vp, vp, vp]
        L.mpres_spmv_csrv.argtypes = [vp, i32, vp, i32, vp, vp, vp]
        L.mpres_lanes_supported.argtypes = [vp, i32]
        L.mpres_set_lanes.argtypes = [vp, i32, i32]
        _lib = L
    return _lib


def check(rc):
    if rc != MPRES_OK:
        raise MpresError(f"mpres error {rc}: {lib().mpres_last_error().decode()}")
    return rc


def set_names():
    L = lib()
    return [L.mpres_moduli_set_name(i).decode() for i in range(L.mpres_num_moduli_sets())]


def _torch():
    import torch

    return torch


class DevBuf:
    """A device byte buffer (torch uint8 tensor) holding numpy data."""

    def __init__(self, nbytes=None, host=None):
        torch = _torch()
        if host is not None:
            host = np.ascontiguousarray(host)
            raw = torch.from_numpy(host.view(np.uint8).reshape(-1).copy())
            self.t = raw.to("cuda")
            self.nbytes = host.nbytes
        else:
            self.nbytes = int(nbytes)
            self.t = torch.zeros(max(self.nbytes, 1), dtype=torch.uint8, device="cuda")

    @property
    def ptr(self):
        return C.c_void_p(self.t.data_ptr())

    def to_numpy(self, dtype, count=None):
        torch = _torch()
        torch.cuda.synchronize()
        b = self.t.cpu().numpy()[: self.nbytes]
        a = b.view(dtype)
        return a if count is None else a[:count]


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

    def set_lanes(self, op, lanes):
        check(self.L.mpres_set_lanes(self.h, op, lanes))

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
        _torch().cuda.synchronize()
        return a

    def elementwise(self, op, x, y, count):
        r = self.aos(count=count)
        check(self.L.mpres_elementwise(self.h, op, count, x.ptr, y.ptr, r.ptr, None))
        return r

    def sync(self):
        _torch().cuda.synchronize()
