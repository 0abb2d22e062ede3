"""ctypes wrapper of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

Numbers are exchanged as numpy structured arrays in the reference's
mp_float_t layout (src/types.cuh:69-74); see ``mp_dtype``.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")

POLICY_E = 0  # reference CPU path (emulated directed rounding)
POLICY_D = 1  # reference CUDA path (IEEE directed rounding)

ER_DTYPE = np.dtype([("frac", "<f8"), ("exp", "<i8")])


def mp_dtype(n):
    """numpy dtype of one mp_float_t with n moduli (4n+40 bytes)."""
    return np.dtype([("digits", "<i4", (n,)), ("sign", "<i4"), ("exp", "<i4"), ("eval", ER_DTYPE, (2,))])


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, ip, dp, i32 = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int
        L.orc_ctx_create.restype = vp
        L.orc_ctx_create.argtypes = [i32]
        L.orc_set_name.restype = C.c_char_p
        for f in ("orc_dadd", "orc_dsub", "orc_dmul", "orc_ddiv"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [i32, i32, C.c_double, C.c_double]
        L.orc_psum.restype = C.c_double
        L.orc_psum.argtypes = [i32, i32, i32, dp]
        L.orc_ceil_log2_lib.argtypes = [C.c_double]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


SET_NAMES = None


def set_names():
    global SET_NAMES
    if SET_NAMES is None:
        L = lib()
        SET_NAMES = [L.orc_set_name(i).decode() for i in range(L.orc_num_sets())]
    return SET_NAMES


class Oracle:
    """One moduli set of the oracle."""

    def __init__(self, set_name):
        L = lib()
        self.L = L
        self.set_name = set_name
        self.ctx = C.c_void_p(L.orc_ctx_create(set_names().index(set_name)))
        self.n = L.orc_n(self.ctx)
        self.log2m = L.orc_log2m(self.ctx)
        self.dtype = mp_dtype(self.n)

    def __del__(self):
        try:
            self.L.orc_ctx_destroy(self.ctx)
        except Exception:
            pass

    # ---- constants ----
    def _ints(self, fn, count):
        out = np.zeros(count, np.int32)
        getattr(self.L, fn)(self.ctx, _p(out))
        return out

    def moduli(self):
        return self._ints("orc_moduli", self.n)

    def part_inverse(self):
        return self._ints("orc_part_inverse", self.n)

    def pow2_table(self):
        return self._ints("orc_pow2_table", (self.log2m + 1) * self.n).reshape(self.log2m + 1, self.n)

    def pow2_residues(self):
        return self._ints("orc_pow2_residues", 30)

    def part_pow2_residues(self):
        return self._ints("orc_part_pow2_residues", 30 * self.n).reshape(30, self.n)

    def pow2_inverse(self):
        return self._ints("orc_pow2_inverse", 30 * self.n).reshape(30, self.n)

    def mrc_inverse(self):
        return self._ints("orc_mrc_inverse", self.n * self.n).reshape(self.n, self.n)

    def recip(self):
        rd = np.zeros(self.n)
        ru = np.zeros(self.n)
        self.L.orc_recip(self.ctx, _p(rd), _p(ru))
        return rd, ru

    def eval_consts(self):
        acc = C.c_double()
        ref = C.c_int()
        er = np.zeros(4, ER_DTYPE)
        self.L.orc_eval_consts(self.ctx, C.byref(acc), C.byref(ref), _p(er[0:1]), _p(er[1:2]), _p(er[2:3]), _p(er[3:4]))
        return dict(acc=acc.value, ref_factor=ref.value, unit_low=er[0], unit_upp=er[1], inv_unit_low=er[2],
                    inv_unit_upp=er[3])

    def mp_consts(self):
        p, h, j = C.c_int(), C.c_int(), C.c_int()
        self.L.orc_mp_consts(self.ctx, C.byref(p), C.byref(h), C.byref(j))
        return p.value, h.value, j.value

    # ---- element ops ----
    def empty(self, count):
        return np.zeros(count, self.dtype)

    def set_d(self, x, policy=POLICY_D):
        x = np.ascontiguousarray(x, np.float64)
        out = self.empty(len(x))
        self.L.orc_set_d(self.ctx, policy, len(x), _p(x), _p(out))
        return out

    def _bin(self, fn, x, y, policy):
        r = self.empty(len(x))
        getattr(self.L, fn)(self.ctx, policy, len(x), _p(np.ascontiguousarray(x)), _p(np.ascontiguousarray(y)), _p(r))
        return r

    def add(self, x, y, policy=POLICY_D):
        return self._bin("orc_add", x, y, policy)

    def sub(self, x, y, policy=POLICY_D):
        return self._bin("orc_sub", x, y, policy)

    def mul(self, x, y, policy=POLICY_D):
        return self._bin("orc_mul", x, y, policy)

    def mul_d(self, x, y, policy=POLICY_D):
        r = self.empty(len(x))
        y = np.ascontiguousarray(y, np.float64)
        self.L.orc_mul_d(self.ctx, policy, len(x), _p(np.ascontiguousarray(x)), _p(y), _p(r))
        return r

    def vadd(self, x, y):
        r = self.empty(len(x))
        self.L.orc_vadd(self.ctx, len(x), _p(np.ascontiguousarray(x)), _p(np.ascontiguousarray(y)), _p(r))
        return r

    def round(self, x, policy=POLICY_D):
        x = np.array(x, copy=True)
        self.L.orc_round(self.ctx, policy, len(x), _p(x))
        return x

    def scale2pow(self, digits, D, policy=POLICY_D):
        digits = np.ascontiguousarray(digits, np.int32)
        D = np.ascontiguousarray(D, np.uint32)
        r = np.zeros_like(digits)
        self.L.orc_scale2pow(self.ctx, policy, len(D), _p(digits), _p(D), _p(r))
        return r

    def eval(self, digits, fast=False, policy=POLICY_D):
        digits = np.ascontiguousarray(digits, np.int32)
        cnt = digits.shape[0]
        lo = np.zeros(cnt, ER_DTYPE)
        up = np.zeros(cnt, ER_DTYPE)
        self.L.orc_eval(self.ctx, policy, int(fast), cnt, _p(digits), _p(lo), _p(up))
        return lo, up

    def rank(self, digits, fast=False, policy=POLICY_D):
        digits = np.ascontiguousarray(digits, np.int32)
        k = np.zeros(digits.shape[0], np.int32)
        self.L.orc_rank(self.ctx, policy, int(fast), digits.shape[0], _p(digits), _p(k))
        return k

    def mrd(self, digits):
        digits = np.ascontiguousarray(digits, np.int32)
        r = np.zeros(digits.shape[0], np.int32)
        self.L.orc_mrd(self.ctx, digits.shape[0], _p(digits), _p(r))
        return r

    def rns_cmp(self, x, y, policy=POLICY_D):
        x = np.ascontiguousarray(x, np.int32)
        y = np.ascontiguousarray(y, np.int32)
        r = np.zeros(x.shape[0], np.int32)
        self.L.orc_rns_cmp(self.ctx, policy, x.shape[0], _p(x), _p(y), _p(r))
        return r

    def cmp(self, x, y, abs_only=False, policy=POLICY_D):
        r = np.zeros(len(x), np.int32)
        self.L.orc_mp_cmp(self.ctx, policy, int(abs_only), len(x), _p(np.ascontiguousarray(x)),
                          _p(np.ascontiguousarray(y)), _p(r))
        return r

    # ---- compositions ----
    def dot(self, x, y, g3, b3, policy=POLICY_D, incx=1, incy=1, n=None):
        n = len(x) if n is None else n
        r = self.empty(1)
        self.L.orc_dot(self.ctx, policy, n, _p(np.ascontiguousarray(x)), incx, _p(np.ascontiguousarray(y)), incy,
                       g3, b3, _p(r))
        return r

    def axpy(self, alpha, x, y, policy=POLICY_D, incx=1, incy=1, n=None):
        n = len(x) if n is None else n
        y = np.array(y, copy=True)
        self.L.orc_axpy(self.ctx, policy, n, _p(np.ascontiguousarray(alpha)), _p(np.ascontiguousarray(x)), incx,
                        _p(y), incy)
        return y

    def axpy_v2(self, alpha, x, y, policy=POLICY_D):
        r = self.empty(len(x))
        self.L.orc_axpy_v2(self.ctx, policy, len(x), _p(np.ascontiguousarray(alpha)), _p(np.ascontiguousarray(x)),
                           _p(np.ascontiguousarray(y)), _p(r))
        return r

    def gemv(self, trans, m, n, alpha, A, lda, x, beta, y, b3, policy=POLICY_D, incx=1, incy=1, rows=None):
        y = np.array(y, copy=True)
        rows_a = None if rows is None else np.ascontiguousarray(rows, np.int32)
        self.L.orc_gemv(self.ctx, policy, int(trans), m, n, _p(np.ascontiguousarray(alpha)),
                        _p(np.ascontiguousarray(A)), lda, _p(np.ascontiguousarray(x)), incx,
                        _p(np.ascontiguousarray(beta)), _p(y), incy, b3, _p(rows_a),
                        0 if rows is None else len(rows_a))
        return y

    def gemm(self, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc, policy=POLICY_D):
        Cm = np.array(Cm, copy=True)
        self.L.orc_gemm(self.ctx, policy, m, n, k, _p(np.ascontiguousarray(alpha)), _p(np.ascontiguousarray(A)), lda,
                        _p(np.ascontiguousarray(B)), ldb, _p(np.ascontiguousarray(beta)), _p(Cm), ldc)
        return Cm

    def gemm_v2(self, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc, policy=POLICY_D, entries=None):
        Cm = np.array(Cm, copy=True)
        e = None if entries is None else np.ascontiguousarray(entries, np.int32).reshape(-1)
        self.L.orc_gemm_v2(self.ctx, policy, int(ta), int(tb), m, n, k, _p(np.ascontiguousarray(alpha)),
                           _p(np.ascontiguousarray(A)), lda, _p(np.ascontiguousarray(B)), ldb,
                           _p(np.ascontiguousarray(beta)), _p(Cm), ldc, _p(e), 0 if e is None else len(e) // 2)
        return Cm

    def add_abs(self, x, y, policy=POLICY_D):
        return self._bin("orc_add_abs", x, y, policy)

    def maxpy_v2(self, alpha, x, y, policy=POLICY_D):
        r = self.empty(len(x))
        self.L.orc_maxpy_v2(self.ctx, policy, len(x), _p(np.ascontiguousarray(alpha)), _p(np.ascontiguousarray(x)),
                            _p(np.ascontiguousarray(y)), _p(r))
        return r

    def sumsq(self, x, blocks, threads, policy=POLICY_D):
        r = self.empty(1)
        self.L.orc_sumsq(self.ctx, policy, len(x), _p(np.ascontiguousarray(x)), blocks, threads, _p(r))
        return r

    # the reference's host MPFR operations (the same under both policies)
    def div(self, x, y):
        r = self.empty(len(x))
        self.L.orc_div(self.ctx, len(x), _p(np.ascontiguousarray(x)), _p(np.ascontiguousarray(y)), _p(r))
        return r

    def sqrt(self, x):
        r = self.empty(len(x))
        self.L.orc_sqrt(self.ctx, len(x), _p(np.ascontiguousarray(x)), _p(r))
        return r

    def dsqrt(self, x):
        out = np.zeros(len(x), np.float64)
        self.L.orc_dsqrt(self.ctx, len(x), _p(np.ascontiguousarray(x)), _p(out))
        return out

    def norm2(self, x, blocks, threads, policy=POLICY_D):
        """cuda::mp_norm2<blocks, threads> (src/blas/v2/norm2_v2.cuh:107-119)"""
        return float(self.dsqrt(self.sumsq(x, blocks, threads, policy))[0])

    def spmv_ell(self, m, maxnzr, as_, ja, x, policy=POLICY_D):
        return self._spmv_fmt(1, m, maxnzr, 0, as_, ja, None, None, None, x, 1, policy)

    def spmv_jad(self, m, maxnzr, as_, ja, jcp, perm, x, tpr=1, policy=POLICY_D):
        return self._spmv_fmt(2, m, maxnzr, 0, as_, ja, jcp, perm, None, x, tpr, policy)

    def spmv_dia(self, m, n, ndiag, as_, offset, x, policy=POLICY_D):
        return self._spmv_fmt(3, m, ndiag, n, as_, None, None, None, offset, x, 1, policy)

    def _spmv_fmt(self, fmt, m, w, ncols, as_, ja, jcp, perm, offset, x, tpr, policy):
        y = self.empty(m)
        arr = [None if a is None else np.ascontiguousarray(a, t) for a, t in
               ((as_, np.float64), (ja, np.int32), (jcp, np.int32), (perm, np.int32), (offset, np.int32))]
        self.L.orc_spmv_fmt(self.ctx, policy, fmt, m, w, ncols, *[_p(a) for a in arr], _p(np.ascontiguousarray(x)),
                            _p(y), tpr)
        return y

    def cg_csr(self, n, irp, ja, as_, b, tol, maxit, x, threads=32, blocks_reduce=256, policy=POLICY_D,
               precond=None, spmv=None):
        """mp_cg_csr<threads, blocks_reduce> (src/sparse/solver/cg_csr.cuh:52-117) or, with precond (the
        reciprocals of A's diagonal), mp_pcg_csr with the diagonal preconditioner (pcg_csr.cuh:52-130),
        composed from the oracle's operations in the reference's order; spmv (v -> A v) replaces the
        CSR product for the solvers of the other formats (cg_ell.cuh, cg_jad.cuh, cg_dia.cuh).
        Returns (x, resvec)."""
        x = np.array(x, copy=True)
        if spmv is None:
            spmv = lambda v: self.spmv(n, irp, ja, as_, v, 1, policy)  # noqa: E731
        r = spmv(x)
        r = self.sub(b, r, policy)
        norm0 = self.norm2(r, blocks_reduce, threads, policy)
        eps = norm0 * tol
        normk = norm0
        resvec = [normk / norm0]
        k = 0
        rho = self.empty(1)
        p = None
        while normk > eps and k < maxit:
            z = r if precond is None else self.mul_d(r, precond, policy)
            rhop = rho
            rho = self.dot(r, z, blocks_reduce, threads, policy)
            if k == 0:
                p = np.array(z, copy=True)
            else:
                beta = self.div(rho, rhop)
                p = self.axpy_v2(beta, p, z, policy)
            q = spmv(p)
            pq = self.dot(p, q, blocks_reduce, threads, policy)
            alpha = self.div(rho, pq)
            x = self.axpy_v2(alpha, p, x, policy)
            r = self.maxpy_v2(alpha, q, r, policy)
            normk = self.norm2(r, blocks_reduce, threads, policy)
            resvec.append(normk / norm0)
            k += 1
        return x, resvec

    def spmv(self, m, irp, ja, as_, x, tpr=1, policy=POLICY_D, rows=None):
        irp = np.ascontiguousarray(irp, np.int32)
        ja = np.ascontiguousarray(ja, np.int32)
        as_ = np.ascontiguousarray(as_, np.float64)
        rows_a = None if rows is None else np.ascontiguousarray(rows, np.int32)
        y = self.empty(m if rows is None else len(rows_a))
        self.L.orc_spmv_csr(self.ctx, policy, m, _p(irp), _p(ja), _p(as_), _p(np.ascontiguousarray(x)), _p(y), tpr,
                            _p(rows_a), 0 if rows is None else len(rows_a))
        return y


def set_threads(n):
    lib().orc_set_threads(int(n))
