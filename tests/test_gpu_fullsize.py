"""GPU parity at the measured shapes, across workgroup and tile boundaries.

Each test runs one entry point at a benchmarked size (or a size that spans
several workgroups / MFMA tiles of every kernel it launches) and compares a
sample of outputs bit for bit with the oracle. The sample reaches every row
workgroup or output tile, the first and last rows, and (for SpMV) the longest
rows. An output of gemv, gemm or SpMV depends only on its own row of A (or
row of op(A) and column of op(B)), so the oracle computes the sampled outputs
on the sub-problem made of those rows and columns: the same arithmetic in the
same order, at a cost of seconds."""
import ctypes as C
import importlib.util
import os
import sys

import numpy as np
import pytest

import exact
import gen
import pyoracle as po

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
D = po.POLICY_D

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from test_gpu_parity import ctx, lanes_of, mfma_reference, mp, note, orc, same, _log2m  # noqa: E402


def tile_sample(size, tile, seed, per_tile=1):
    """first and last index of every tile of `tile` indices, plus `per_tile` random ones inside"""
    rng = np.random.default_rng(seed)
    out = set()
    for t0 in range(0, size, tile):
        t1 = min(size, t0 + tile) - 1
        out.update((t0, t1))
        out.update(int(v) for v in rng.integers(t0, t1 + 1, per_tile))
    return np.array(sorted(out), np.int64)


def colmajor_rows(vals, m, n, rows):
    """the rows `rows` of an m x n column-major matrix, as a len(rows) x n column-major matrix"""
    return np.ascontiguousarray(vals.reshape(n, m)[:, rows]).ravel()


def colmajor_cols(vals, m, n, cols):
    return np.ascontiguousarray(vals.reshape(n, m)[cols, :]).ravel()


def test_gemv_4096_32moduli():
    """mp_gemv 4096 x 4096 at 32 moduli (424 bits), e in [-30, 30], block3 32: three rows of
    every 128-row workgroup of the residue kernel (first, last, one inside), both
    transposes"""
    name, m, n, b3 = "32_8double", 4096, 4096, 32
    c, o = ctx(name), orc(name)
    Ad = gen.uexp(201, m * n, -30, 30)
    xd, yd = gen.uexp(202, n, -30, 30), gen.uexp(203, m, -30, 30)
    al, be = o.set_d(gen.uexp(204, 1, -30, 30), D), o.set_d(gen.uexp(205, 1, -30, 30), D)
    aA = c.array_set_d(Ad)
    aa, ab = c.array_from_host(al), c.array_from_host(be)
    rows = tile_sample(m, 128, 206)
    for trans in (0, 1):
        note("gemv 4096^2 32 moduli trans", trans, "rows", len(rows))
        ax, ay = c.array_set_d(xd), c.array_set_d(yd)
        c.set_lanes(mp.K_GEMV, 0)
        mp.check(c.L.mpres_gemv(c.h, mp.TRANS if trans else mp.NO_TRANS, m, n, C.byref(aa), C.byref(aA), m,
                                C.byref(ax), 1, C.byref(ab), C.byref(ay), 1, b3, None, 0, None))
        got = c.array_to_host(ay, m)[rows]
        # output i uses row i of op(A): column i of A when transposed
        sub = colmajor_cols(Ad, m, n, rows).reshape(len(rows), n).T.ravel() if trans else \
            colmajor_rows(Ad, m, n, rows)
        want = o.gemv(0, len(rows), n, al, o.set_d(sub, D), len(rows), o.set_d(xd, D), be, o.set_d(yd[rows], D),
                      b3, D)
        same(got, want, f"gemv 4096 trans={trans}")


def test_gemm_512_32moduli():
    """mp_gemm 512^3 at 32 moduli, e in [-30, 30]: two rows of every 64-row tile times two
    columns of every 64-column tile, so every 64 x 64 tile of the INT8 residue products and
    every block of the chain and digit kernels is sampled"""
    name, m, n, k = "32_8double", 512, 512, 512
    c, o = ctx(name), orc(name)
    Ad, Bd, Cd = gen.uexp(211, m * k, -30, 30), gen.uexp(212, k * n, -30, 30), gen.uexp(213, m * n, -30, 30)
    al, be = o.set_d(gen.uexp(214, 1, -30, 30), D), o.set_d(gen.uexp(215, 1, -30, 30), D)
    aA, aB, aC = c.array_set_d(Ad), c.array_set_d(Bd), c.array_set_d(Cd)
    aa, ab = c.array_from_host(al), c.array_from_host(be)
    c.set_lanes(mp.K_GEMM, 0)
    note("gemm 512^3 32 moduli")
    mp.check(c.L.mpres_gemm(c.h, mp.NO_TRANS, mp.NO_TRANS, m, n, k, C.byref(aa), C.byref(aA), m, C.byref(aB), k,
                            C.byref(ab), C.byref(aC), m, None, 0, None))
    got = c.array_to_host(aC, m * n).reshape(n, m)
    rows, cols = tile_sample(m, 64, 216, 0), tile_sample(n, 64, 217, 0)
    As = o.set_d(colmajor_rows(Ad, m, k, rows), D)
    Bs = o.set_d(colmajor_cols(Bd, k, n, cols), D)
    Cs = o.set_d(Cd.reshape(n, m)[np.ix_(cols, rows)].ravel(), D)
    want = o.gemm(len(rows), len(cols), k, al, As, len(rows), Bs, k, be, Cs, len(rows), D)
    same(got[np.ix_(cols, rows)].ravel(), want, "gemm 512^3")


@pytest.mark.parametrize("name", ["8_2double", "16_4double", "32_8double"])
def test_gemm_multitile(name):
    """v1 and v2 gemm at 130 x 70 x 40 (several 64 x 64 tiles and 256-entry blocks, partial
    edge tiles) for every lanes variant, all entries against the oracle; v2 with every
    transpose pair"""
    c, o = ctx(name), orc(name)
    m, n, k = 130, 70, 40
    A = o.set_d(gen.uexp(221, m * k, -30, 30), D)
    B = o.set_d(gen.uexp(222, k * n, -30, 30), D)
    Cm = o.set_d(gen.uexp(223, m * n, -30, 30), D)
    alpha, beta = o.set_d(gen.uexp(224, 1, -30, 30), D), o.set_d(gen.uexp(225, 1, -30, 30), D)
    want = o.gemm(m, n, k, alpha, A, m, B, k, beta, Cm, m, D)
    wants = {(ta, tb): o.gemm_v2(ta, tb, m, n, k, alpha, A, k if ta else m, B, n if tb else k, beta, Cm, m, D)
             for ta in (0, 1) for tb in (0, 1)}
    aa, ab = c.array_from_host(alpha), c.array_from_host(beta)
    aA, aB = c.array_from_host(A), c.array_from_host(B)
    dA, dB, da, db = c.aos(A), c.aos(B), c.aos(alpha), c.aos(beta)
    for L in lanes_of(c):
        c.set_lanes(mp.K_GEMM, L)
        note("gemm multitile", name, m, n, k, L)
        aC = c.array_from_host(Cm)
        mp.check(c.L.mpres_gemm(c.h, mp.NO_TRANS, mp.NO_TRANS, m, n, k, C.byref(aa), C.byref(aA), m, C.byref(aB), k,
                                C.byref(ab), C.byref(aC), m, None, 0, None))
        same(c.array_to_host(aC, m * n), want, f"gemm {name} L={L}")
        for (ta, tb), w in wants.items():
            dC = c.aos(Cm)
            mp.check(c.L.mpres_gemm_v2(c.h, mp.TRANS if ta else mp.NO_TRANS, mp.TRANS if tb else mp.NO_TRANS, m, n,
                                       k, da.ptr, dA.ptr, k if ta else m, dB.ptr, n if tb else k, db.ptr, dC.ptr, m,
                                       None, 0, None))
            same(c.aos_to_host(dC, m * n), w, f"gemm_v2 {name} {ta}{tb} L={L}")


@pytest.mark.parametrize("name,b3", [("32_8double", 8), ("16_4double", 32), ("8_2double", 8)])
def test_gemv_multiblock(name, b3):
    """gemv 300 x 260 (three row workgroups of the residue kernel, a partial last one) for
    every lanes variant and both transposes, all outputs against the oracle"""
    c, o = ctx(name), orc(name)
    m, n = 300, 260
    A = o.set_d(gen.uexp(231, m * n, -30, 30), D)
    alpha, beta = o.set_d(gen.uexp(232, 1, -30, 30), D), o.set_d(gen.uexp(233, 1, -30, 30), D)
    aA, aa, ab = c.array_from_host(A), c.array_from_host(alpha), c.array_from_host(beta)
    for trans in (0, 1):
        lx, ly = (m, n) if trans else (n, m)
        x = o.set_d(gen.uexp(234, lx, -30, 30), D)
        y = o.set_d(gen.uexp(235, ly, -30, 30), D)
        want = o.gemv(trans, m, n, alpha, A, m, x, beta, y, b3, D)
        ax = c.array_from_host(x)
        for L in lanes_of(c):
            c.set_lanes(mp.K_GEMV, L)
            note("gemv multiblock", name, m, n, b3, trans, L)
            ay = c.array_from_host(y)
            mp.check(c.L.mpres_gemv(c.h, mp.TRANS if trans else mp.NO_TRANS, m, n, C.byref(aa), C.byref(aA), m,
                                    C.byref(ax), 1, C.byref(ab), C.byref(ay), 1, b3, None, 0, None))
            same(c.array_to_host(ay, ly), want, f"gemv {name} trans={trans} L={L}")


@pytest.mark.parametrize("name,lo,hi", [("8_2double", -8, 8), ("8_2double", -60, 60), ("32_8double", -8, 8),
                                        ("32_8double", -60, 60)])
def test_gemm_mfma_multitile(name, lo, hi):
    """precision form at 300 x 260 x 200 (five by five 64 x 64 MFMA tiles, partial edge
    tiles, k over several 32-deep k-steps), 8 and 32 moduli, narrow and wide exponents:
    two rows of every row tile times two columns of every column tile, against the
    exact-integer restatement of the form"""
    c, o = ctx(name), orc(name)
    m, n, k = 300, 260, 200
    Ad, Bd, Cd = gen.uexp(241, m * k, lo, hi), gen.uexp(242, k * n, lo, hi), gen.uexp(243, m * n, lo, hi)
    alpha, beta = o.set_d([1.25], D), o.set_d([-0.5], D)
    aA, aB, aC = c.array_set_d(Ad), c.array_set_d(Bd), c.array_set_d(Cd)
    aa, ab = c.array_from_host(alpha), c.array_from_host(beta)
    c.set_lanes(mp.K_MFMA, 0)
    note("gemm_mfma multitile", name, lo, hi)
    mp.check(c.L.mpres_gemm_mfma(c.h, mp.NO_TRANS, mp.NO_TRANS, m, n, k, C.byref(aa), C.byref(aA), m, C.byref(aB), k,
                                 C.byref(ab), C.byref(aC), m, None, 0, None))
    got = c.array_to_host(aC, m * n).reshape(n, m)
    rows, cols = tile_sample(m, 64, 244, 0), tile_sample(n, 64, 245, 0)
    As = o.set_d(colmajor_rows(Ad, m, k, rows), D)
    Bs = o.set_d(colmajor_cols(Bd, k, n, cols), D)
    Cs = o.set_d(Cd.reshape(n, m)[np.ix_(cols, rows)].ravel(), D)
    mr, nc = len(rows), len(cols)
    want, _ = mfma_reference(o, _log2m(c), 0, 0, mr, nc, k, alpha, As, mr, Bs, k, beta, Cs, mr)
    same(got[np.ix_(cols, rows)].ravel(), want, f"gemm_mfma {name} [{lo},{hi}]")


_SPMV = {}


def _spmv_headline():
    if not _SPMV:
        m = 1_000_000
        irp, ja, as_ = gen.powerlaw_csr(7, m)
        _SPMV.update(m=m, irp=irp, ja=ja, as_=as_, x=gen.uexp(8, m, -10, 10))
    return _SPMV


@pytest.mark.parametrize("tpr", [1, 4, 16, 32])
def test_spmv_1e6_16moduli(tpr):
    """SpMV on the 10^6-row power-law matrix (nnz per row 1..1024, mean 16) at 16 moduli,
    x with e in [-10, 10]: about 600 rows spread over the whole matrix, the first and last
    rows and every row of 1000 or more nonzeros, against the oracle"""
    name = "16_4double"
    c, o = ctx(name), orc(name)
    S = _spmv_headline()
    m, irp, ja, as_ = S["m"], S["irp"], S["ja"], S["as_"]
    lens = np.diff(irp)
    rows = set(tile_sample(m, 2048, 250 + tpr, 0).tolist()) | set(np.nonzero(lens >= 1000)[0].tolist())
    rows = np.array(sorted(rows), np.int32)
    assert lens[rows].max() >= 1000
    dirp, dja, das = mp.DevBuf(host=irp), mp.DevBuf(host=ja), mp.DevBuf(host=as_)
    csr = mp.Csr(das.ptr.value, dja.ptr.value, dirp.ptr.value)
    dx = c.set_d(S["x"])
    dy = c.aos(count=m)
    c.set_lanes(mp.K_SPMV, 0)
    note("spmv 1e6 16 moduli tpr", tpr, "rows", len(rows), "nnz", int(irp[-1]))
    if tpr == 1:
        mp.check(c.L.mpres_spmv_csr(c.h, m, C.byref(csr), dx.ptr, dy.ptr, None))
    else:
        mp.check(c.L.mpres_spmv_csrv(c.h, m, C.byref(csr), tpr, dx.ptr, dy.ptr, None))
    got = c.aos_to_host(dy, m)[rows]
    # only the x entries these rows read are needed; the oracle reads x by column index
    x = np.zeros(m, o.dtype)
    cols = np.unique(np.concatenate([ja[irp[r]:irp[r + 1]] for r in rows]))
    x[cols] = o.set_d(S["x"][cols], D)
    want = o.spmv(m, irp, ja, as_, x, tpr, D, rows=rows)
    same(got, want, f"spmv 1e6 tpr={tpr}")
    # exact: at 212 bits no row sum rounds; compare a few rows with the exact sum
    mm = [int(v) for v in o.moduli()]
    for r in rows[:: max(1, len(rows) // 8)]:
        from fractions import Fraction
        ex = sum(Fraction(float(as_[i])) * Fraction(float(S["x"][ja[i]])) for i in range(irp[r], irp[r + 1]))
        assert exact.mp_value(want[list(rows).index(r)], mm) == ex
