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

    # ---- the rest of the BLAS (src/blas, src/blas/v2): compositions of the element operations
    # above in the reference's kernel order. Numbers are AoS records here whatever layout the
    # library keeps them in; x[bidx] is element i of a vector with increment inc.
    def _bcast(self, a, count):
        return np.repeat(np.ascontiguousarray(a)[:1], count)

    def _vsub(self, x, y):
        """mp_vector_sub_esi_kernel + mp_vector_add_digits_kernel (src/mpvector.cuh:487-560):
        the v1 add with y's sign inverted"""
        y = np.array(y, copy=True)
        y["sign"] ^= 1
        return self.vadd(x, y)

    def scal(self, n, alpha, x, incx):
        """x = alpha x: cuda::mp_scal (src/blas/scal.cuh:43-61)"""
        x = np.array(x, copy=True)
        if n <= 0 or incx <= 0:
            return x
        i = bidx(np.arange(n), n, incx)
        x[i] = self.mul(x[i], self._bcast(alpha, n))
        return x

    def rot(self, n, x, incx, y, incy, c, s):
        """cuda::mp_rot (src/blas/rot.cuh:49-100): returns (x, y, buffer1, buffer2); the c scaling
        runs with increment 1 (:78-81)"""
        x, y = np.array(x, copy=True), np.array(y, copy=True)
        ix, iy = bidx(np.arange(n), n, incx), bidx(np.arange(n), n, incy)
        S, Cc = self._bcast(s, n), self._bcast(c, n)
        b1 = self.mul(x[ix], S)
        b2 = self.mul(y[iy], S)
        x[:n] = self.mul(x[:n], Cc)
        y[:n] = self.mul(y[:n], Cc)
        x[ix] = self.vadd(x[ix], b2)
        y[iy] = self._vsub(y[iy], b1)
        return x, y, b1, b2

    def waxpby(self, n, alpha, x, incx, beta, y, incy, w, incw):
        """w = alpha x + beta y: cuda::mp_waxpby (src/blas/waxpby.cuh:51-89); returns (w, buffer)"""
        w = np.array(w, copy=True)
        ar = np.arange(n)
        buf = self.mul(x[bidx(ar, n, incx)], self._bcast(alpha, n))
        iw = bidx(ar, n, incw)
        w[iw] = self.mul(y[bidx(ar, n, incy)], self._bcast(beta, n))
        w[iw] = self.vadd(w[iw], buf)
        return w, buf

    def axpy_dot(self, n, alpha, w, incw, v, incv, u, incu, g3, b3):
        """w = w - alpha v, r = u . w: cuda::mp_axpy_dot (src/blas/axpydot.cuh:54-86); returns (w, r)"""
        w = np.array(w, copy=True)
        ar = np.arange(n)
        buf = self.mul(v[bidx(ar, n, incv)], self._bcast(alpha, n))
        iw = bidx(ar, n, incw)
        w[iw] = self._vsub(w[iw], buf)
        return w, self.dot(u, w, g3, b3, POLICY_D, incu, incw, n)

    def _fold_tree(self, terms_at, cnt, g, b, step):
        """the reduction kernels' shape (src/mpreduct.cuh:120-286): virtual thread v of g*b folds
        terms v, v + g b, ... from zero with step(acc, term); then, per block of b, the halving tree
        from the least power of two >= b. Returns the g block results."""
        G = g * b
        acc = self.empty(G)
        for k in range((cnt + G - 1) // G):
            v = np.arange(G)
            idx = v + k * G
            m = idx < cnt
            acc[m] = step(acc[m], terms_at(idx[m]))
        acc = acc.reshape(g, b)
        h = next_pow2(b) >> 1
        while h >= 1:
            t = np.arange(min(h, b))
            t = t[t + h < b]
            if len(t):
                lhs = acc[:, t].reshape(-1)
                rhs = acc[:, t + h].reshape(-1)
                acc[:, t] = step(lhs, rhs).reshape(g, len(t))
            h >>= 1
        return np.ascontiguousarray(acc[:, 0])

    def _max_step(self, acc, t):
        """keep t where |t| > |acc| (mp_cmp_abs == 1)"""
        r = np.array(acc, copy=True)
        sel = self.cmp(t, acc, abs_only=True) == 1
        r[sel] = t[sel]
        return r

    def asum(self, n, x, incx, g1, b1):
        """r = sum |x_i|: cuda::mp_asum<g1, b1> (src/blas/asum.cuh:41-70), with
        mp_array_reduce_sum_abs_kernel1/2 (src/mpreduct.cuh:120-192); mp_add_abs throughout"""
        add = lambda a, t: self.add_abs(a, t)  # noqa: E731
        p = self._fold_tree(lambda i: x[i * incx], n, g1, b1, add)
        r = self._fold_tree(lambda i: p[i], g1, 1, b1, add)
        r["sign"] = 0
        return r

    def norm_inf(self, n, x, incx, g1, b1):
        """r = max |x_i|: cuda::mp_norm<g1, b1>(mblas_inf_norm) (src/blas/norm.cuh:43-79), with
        mp_array_reduce_max_abs_kernel1/2 (src/mpreduct.cuh:199-286)"""
        p = self._fold_tree(lambda i: x[i * incx], n, g1, b1, self._max_step)
        p["sign"] = 0
        r = self._fold_tree(lambda i: p[i], g1, 1, b1, self._max_step)
        r["sign"] = 0
        return r

    def _ge_map(self, m, n, A, lda, f):
        """indices of the leading m x n part of a column-major array with leading dimension lda"""
        i, j = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
        return (i + j * lda).reshape(-1, order="F")

    def ge_add(self, m, n, alpha, A, lda, beta, B, ldb, Cm, ldc):
        """C = alpha A + beta B: cuda::mp_ge_add (src/blas/geadd.cuh:58-114); returns (C, buffer)"""
        Cm = np.array(Cm, copy=True)
        ia, ib, ic = (self._ge_map(m, n, None, ld, None) for ld in (lda, ldb, ldc))
        buf = self.mul(A[ia], self._bcast(alpha, m * n))
        Cm[ic] = self.mul(B[ib], self._bcast(beta, m * n))
        Cm[ic] = self.vadd(Cm[ic], buf)
        return Cm, buf

    def ge_acc(self, m, n, alpha, A, lda, beta, B, ldb):
        """B = alpha A + beta B: cuda::mp_ge_acc (src/blas/geacc.cuh:57-113); returns (B, buffer)"""
        B = np.array(B, copy=True)
        ia, ib = (self._ge_map(m, n, None, ld, None) for ld in (lda, ldb))
        buf = self.mul(A[ia], self._bcast(alpha, m * n))
        B[ib] = self.mul(B[ib], self._bcast(beta, m * n))
        B[ib] = self.vadd(B[ib], buf)
        return B, buf

    def ge_diag_scale(self, side_right, m, n, D, incd, A, lda):
        """A = A D (right) or D A (left): cuda::mp_ge_diag_scale (src/blas/gediagscale.cuh:54-91),
        the factor of element (i, j) D[bidx(j, n, incd)] or D[bidx(i, m, incd)]
        (mp_mat2vec_right/left_scal_esi_kernel, src/mpmatrix.cuh:259-404)"""
        A = np.array(A, copy=True)
        ia = self._ge_map(m, n, None, lda, None)
        i, j = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
        i, j = i.reshape(-1, order="F"), j.reshape(-1, order="F")
        d = D[bidx(j, n, incd)] if side_right else D[bidx(i, m, incd)]
        A[ia] = self.mul(A[ia], d)
        return A

    def ge_lr_scale(self, m, n, DL, incdl, DR, incdr, A, lda):
        """A = DL A DR: cuda::mp_ge_lr_scale (src/blas/gelrscale.cuh:56-93), left first"""
        A = self.ge_diag_scale(False, m, n, DL, incdl, A, lda)
        return self.ge_diag_scale(True, m, n, DR, incdr, A, lda)

    def ge_norm(self, one_norm, m, n, A, lda, g1, b1):
        """cuda::mp_ge_norm<g1, b1> (src/blas/genorm.cuh:142-173): per column (one-norm) or row
        (infinity-norm) the sum of |a| by b1 threads striding over it from zero and the halving
        tree (matrix_col/row_sum_abs_kernel, :38-139), written with sign 0; then mp_norm's
        infinity norm of those sums. Returns (r, sums)."""
        outs, ln = (n, m) if one_norm else (m, n)
        add = lambda a, t: self.add_abs(a, t)  # noqa: E731
        acc = self.empty(outs * b1).reshape(outs, b1)
        o = np.arange(outs)
        for k in range((ln + b1 - 1) // b1):
            for t in range(b1):
                i = t + k * b1
                if i >= ln:
                    continue
                idx = (i + o * lda) if one_norm else (o + i * lda)
                acc[:, t] = add(acc[:, t], A[idx])
        h = next_pow2(b1) >> 1
        while h >= 1:
            for t in range(min(h, b1)):
                if t + h < b1:
                    acc[:, t] = add(acc[:, t], acc[:, t + h])
            h >>= 1
        sums = np.ascontiguousarray(acc[:, 0])
        sums["sign"] = 0
        return self.norm_inf(outs, sums, 1, g1, b1), sums

    def ger(self, m, n, alpha, x, incx, y, incy, A, lda):
        """A = A + x (alpha y)^T: cuda::mp_ger (src/blas/ger.cuh:155-214); returns (A, buffer1,
        buffer2)"""
        A = np.array(A, copy=True)
        b1 = self.mul(y[bidx(np.arange(n), n, incy)], self._bcast(alpha, n))
        i, j = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
        i, j = i.reshape(-1, order="F"), j.reshape(-1, order="F")
        b2 = self.mul(x[bidx(i, m, incx)], b1[j])  # mp_cvec2rvec_mul_esi/digits_kernel (:39-150)
        ia = i + j * lda
        A[ia] = self.vadd(A[ia], b2)
        return A, b1, b2

    # v2 (mp_float_t arrays)
    def scal_v2(self, alpha, x):
        """cuda::mp_scal (src/blas/v2/scal_v2.cuh:37-48): r = alpha x"""
        return self.mul(self._bcast(alpha, len(x)), x)

    def ger_v2(self, m, n, alpha, x, incx, y, incy, A, lda):
        """cuda::mp_ger (src/blas/v2/ger_v2.cuh:49-68): A += (alpha x_i) y_j"""
        A = np.array(A, copy=True)
        i, j = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
        i, j = i.reshape(-1, order="F"), j.reshape(-1, order="F")
        t = self.mul(self._bcast(alpha, m * n), x[bidx(i, m, incx)])
        t = self.mul(t, y[bidx(j, n, incy)])
        ia = i + j * lda
        A[ia] = self.add(A[ia], t)
        return A

    def _tri(self, upper, n):
        i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        i, j = i.reshape(-1, order="F"), j.reshape(-1, order="F")
        keep = (i <= j) if upper else (j <= i)
        return i[keep], j[keep]

    def syr_v2(self, upper, n, alpha, x, incx, A, lda):
        """cuda::mp_syr (src/blas/v2/syr_v2.cuh:58-76): A += (x_i x_j) alpha on the triangle"""
        A = np.array(A, copy=True)
        i, j = self._tri(upper, n)
        t = self.mul(x[bidx(i, n, incx)], x[bidx(j, n, incx)])
        t = self.mul(t, self._bcast(alpha, len(i)))
        ia = i + j * lda
        A[ia] = self.add(A[ia], t)
        return A

    def syr2_v2(self, upper, n, alpha, x, incx, y, incy, A, lda):
        """cuda::mp_syr2 (src/blas/v2/syr2_v2.cuh:59-83): A += (x_i y_j) alpha + (y_i x_j) alpha"""
        A = np.array(A, copy=True)
        i, j = self._tri(upper, n)
        a = self._bcast(alpha, len(i))
        t1 = self.mul(self.mul(x[bidx(i, n, incx)], y[bidx(j, n, incy)]), a)
        t2 = self.mul(self.mul(y[bidx(i, n, incy)], x[bidx(j, n, incx)]), a)
        ia = i + j * lda
        A[ia] = self.add(A[ia], self.add(t1, t2))
        return A

    def symv_v2(self, upper, n, alpha, A, lda, x, incx, beta, y, incy):
        """cuda::mp_symv (src/blas/v2/symv_v2.cuh:57-95): per row, sum over columns in order of
        x_col A(row, col) from the stored triangle, then y = (beta y) + (alpha sum)"""
        y = np.array(y, copy=True)
        rows = np.arange(n)
        s = self.empty(n)
        for col in range(n):
            lower_part = rows <= col
            if not upper:
                lower_part = ~lower_part
            ia = np.where(lower_part, rows + col * lda, col + rows * lda)
            s = self.add(s, self.mul(self._bcast(x[bidx(col, n, incx):], n), A[ia]))
        iy = bidx(rows, n, incy)
        yb = self.mul(self._bcast(beta, n), y[iy])
        s = self.mul(self._bcast(alpha, n), s)
        y[iy] = self.add(yb, s)
        return y

    def syrk_v2(self, upper, trans, n, k, alpha, A, lda, beta, Cm, ldc):
        """cuda::mp_syrk (src/blas/v2/syrk_v2.cuh:61-85): C(row, col) = (sum_i op(A)(row, i)
        op(A)(col, i)) alpha + C(row, col) beta on the triangle; only trans reads A transposed"""
        Cm = np.array(Cm, copy=True)
        r, c = self._tri(upper, n)
        s = self.empty(len(r))
        for i in range(k):
            ia = (i + lda * r) if trans else (r + lda * i)
            ib = (i + lda * c) if trans else (c + lda * i)
            s = self.add(s, self.mul(A[ia], A[ib]))
        s = self.mul(s, self._bcast(alpha, len(r)))
        ic = r + c * ldc
        t = self.mul(Cm[ic], self._bcast(beta, len(r)))
        Cm[ic] = self.add(s, t)
        return Cm


def bidx(i, n, inc):
    """index of element i of a vector of n with increment inc (negative: from the end)"""
    i = np.asarray(i, np.int64)
    return i * inc if inc > 0 else (-n + i + 1) * inc


def next_pow2(v):
    p = 1
    while p < v:
        p <<= 1
    return p


def set_threads(n):
    lib().orc_set_threads(int(n))
