"""GPU parity of the rest of the BLAS (mpres_scal, rot, waxpby, axpy_dot, asum, norm, ge_add,
ge_acc, ge_diag_scale, ge_lr_scale, ge_norm, ger and the v2 scal, prod, ger, syr, syr2, symv,
syrk) against the oracle's restatement of the reference kernels (policy D), bit for bit in
every field, on operands with full-length significands, cancellations, zeros and negative
increments, for every compiled lanes-per-number variant."""
import ctypes as C
import importlib.util
import os
import sys

import numpy as np
import pytest

import gen
import pyoracle as po

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
D = po.POLICY_D


def _load_mpres():
    spec = importlib.util.spec_from_file_location("mpres", os.path.join(ROOT, "mpres-blas_amd", "mpres.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


mp = _load_mpres()
_ctx, _orc = {}, {}
SETS = ["8_2double", "16_4double", "32_8double"]


def ctx(name):
    if name not in _ctx:
        _ctx[name] = mp.Context(name)
    return _ctx[name]


def orc(name):
    if name not in _orc:
        _orc[name] = po.Oracle(name)
    return _orc[name]


def lanes_of(c):
    return [L for L in (1, 2, 4, 8, 16, 32) if c.L.mpres_lanes_supported(c.h, L)]


def note(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def same(got, want, what):
    got, want = np.ascontiguousarray(got), np.ascontiguousarray(want)
    assert got.dtype == want.dtype and got.shape == want.shape
    if got.tobytes() == want.tobytes():
        return
    bad = [i for i in range(len(got)) if got[i].tobytes() != want[i].tobytes()]
    i = bad[0]
    raise AssertionError(f"{what}: {len(bad)} of {len(got)} records differ; first {i}:\n gpu {got[i]}\n ref {want[i]}")


def vals(o, seed, count, elo=-30, ehi=30):
    """converted doubles (some zero), products with full-length significands, and negated
    near-copies of earlier entries (cancellation in the adds)"""
    a = o.set_d(gen.uexp(seed, count, elo, ehi), D)
    b = o.set_d(gen.uexp(seed + 1, count, elo, ehi), D)
    p = o.mul(a, b, D)
    r = np.where(np.arange(count) % 3 == 0, a, p)
    r[::11] = o.set_d(np.zeros(len(r[::11])), D)
    return r


def scalar(o, v):
    return o.set_d([v], D)


@pytest.mark.parametrize("name", SETS)
def test_v1_vector_routines(name):
    c, o = ctx(name), orc(name)
    n = 500
    x0, y0, w0 = vals(o, 1, 3 * n), vals(o, 3, 3 * n), vals(o, 5, 3 * n)
    al, be = o.mul(vals(o, 7, 1), scalar(o, -1.5), D), scalar(o, 0.375)
    cc, ss = vals(o, 9, 2)[1:2], o.mul(vals(o, 9, 2)[1:2], scalar(o, 0.5), D)
    u0, v0 = vals(o, 11, 2 * n), vals(o, 13, 2 * n)
    want_scal = o.scal(n, al, x0, 3)
    rot_cases = [(1, 1), (2, -3)]
    want_rot = [o.rot(n, x0, ix, y0, iy, cc, ss) for ix, iy in rot_cases]
    want_wax = o.waxpby(n, al, x0, -2, be, y0, 1, w0, 3)
    want_ad = o.axpy_dot(n, al, w0, 2, v0, -1, u0, 2, 16, 32)
    for L in lanes_of(c):
        c.set_lanes(mp.K_ELEM, L)
        c.set_lanes(mp.K_DOT, L)
        note("v1 vector routines", name, L)
        A = c.array_from_host
        ax, aal = A(x0), A(al)
        mp.check(c.L.mpres_scal(c.h, n, C.byref(aal), C.byref(ax), 3, None))
        same(c.array_to_host(ax, len(x0)), want_scal, f"scal {name} L={L}")
        for (ix, iy), (wx, wy, wb1, wb2) in zip(rot_cases, want_rot):
            ax, ay, acc, ass, b1, b2 = A(x0), A(y0), A(cc), A(ss), c.array(n), c.array(n)
            mp.check(c.L.mpres_rot(c.h, n, C.byref(ax), ix, C.byref(ay), iy, C.byref(acc), C.byref(ass), C.byref(b1),
                                   C.byref(b2), None))
            same(c.array_to_host(ax, len(x0)), wx, f"rot x {name} L={L} inc {ix},{iy}")
            same(c.array_to_host(ay, len(y0)), wy, f"rot y {name} L={L} inc {ix},{iy}")
            same(c.array_to_host(b1, n), wb1, f"rot buffer1 {name} L={L}")
            same(c.array_to_host(b2, n), wb2, f"rot buffer2 {name} L={L}")
        ax, ay, aw, abe, buf = A(x0), A(y0), A(w0), A(be), c.array(n)
        mp.check(c.L.mpres_waxpby(c.h, n, C.byref(aal), C.byref(ax), -2, C.byref(abe), C.byref(ay), 1, C.byref(aw), 3,
                                  C.byref(buf), None))
        same(c.array_to_host(aw, len(w0)), want_wax[0], f"waxpby {name} L={L}")
        same(c.array_to_host(buf, n), want_wax[1], f"waxpby buffer {name} L={L}")
        aw, av, au, ar, buf = A(w0), A(v0), A(u0), c.array(1), c.array(n)
        mp.check(c.L.mpres_axpy_dot(c.h, n, C.byref(aal), C.byref(aw), 2, C.byref(av), -1, C.byref(au), 2,
                                    C.byref(ar), C.byref(buf), 16, 32, None, 0, None))
        same(c.array_to_host(aw, len(w0)), want_ad[0], f"axpy_dot w {name} L={L}")
        same(c.array_to_host(ar, 1), want_ad[1], f"axpy_dot r {name} L={L}")


@pytest.mark.parametrize("name", SETS)
@pytest.mark.parametrize("n,incx,g1,b1", [(3000, 1, 32, 64), (777, 3, 5, 24), (50, 2, 128, 128), (1, 1, 1, 1)])
def test_asum_norm(name, n, incx, g1, b1):
    c, o = ctx(name), orc(name)
    x = vals(o, 21, n * incx)
    x["sign"][::2] ^= 1
    want_sum = o.asum(n, x, incx, g1, b1)
    want_max = o.norm_inf(n, x, incx, g1, b1)
    for L in lanes_of(c):
        c.set_lanes(mp.K_DOT, L)
        note("asum / norm", name, n, incx, g1, b1, L)
        ax, ar = c.array_from_host(x), c.array(1)
        mp.check(c.L.mpres_asum(c.h, n, C.byref(ax), incx, C.byref(ar), g1, b1, None, 0, None))
        same(c.array_to_host(ar, 1), want_sum, f"asum {name} L={L}")
        mp.check(c.L.mpres_norm(c.h, mp.ONE_NORM, n, C.byref(ax), incx, C.byref(ar), g1, b1, None, 0, None))
        same(c.array_to_host(ar, 1), want_sum, f"norm one {name} L={L}")
        mp.check(c.L.mpres_norm(c.h, mp.INF_NORM, n, C.byref(ax), incx, C.byref(ar), g1, b1, None, 0, None))
        same(c.array_to_host(ar, 1), want_max, f"norm inf {name} L={L}")


@pytest.mark.parametrize("name", SETS)
def test_ge_routines(name):
    c, o = ctx(name), orc(name)
    m, n, lda, ldb, ldc = 37, 23, 40, 37, 41
    A0, B0, C0 = vals(o, 31, lda * n), vals(o, 33, ldb * n), vals(o, 35, ldc * n)
    al, be = vals(o, 37, 2)[1:2], scalar(o, -0.75)
    D0 = vals(o, 39, 2 * max(m, n))
    want_add = o.ge_add(m, n, al, A0, lda, be, B0, ldb, C0, ldc)
    want_acc = o.ge_acc(m, n, al, A0, lda, be, B0, ldb)
    diag_cases = [(mp.RIGHT, 2), (mp.LEFT, -1)]
    want_diag = [o.ge_diag_scale(s == mp.RIGHT, m, n, D0, inc, A0, lda) for s, inc in diag_cases]
    want_lr = o.ge_lr_scale(m, n, D0, -2, D0, 1, A0, lda)
    norm_cases = [(mp.ONE_NORM, 4, 16), (mp.INF_NORM, 3, 7)]
    want_norm = [o.ge_norm(nm == mp.ONE_NORM, m, n, A0, lda, g, b) for nm, g, b in norm_cases]
    x0, y0 = vals(o, 41, 2 * m), vals(o, 43, 3 * n)
    ger_cases = [(1, 1), (-2, 3)]
    want_ger = [o.ger(m, n, al, x0, ix, y0, iy, A0, lda) for ix, iy in ger_cases]
    for L in lanes_of(c):
        c.set_lanes(mp.K_ELEM, L)
        c.set_lanes(mp.K_DOT, L)
        note("ge routines", name, L)
        A = c.array_from_host
        aA, aB, aC, aal, abe, buf = A(A0), A(B0), A(C0), A(al), A(be), c.array(m * n)
        mp.check(c.L.mpres_ge_add(c.h, m, n, C.byref(aal), C.byref(aA), lda, C.byref(abe), C.byref(aB), ldb,
                                  C.byref(aC), ldc, C.byref(buf), None))
        same(c.array_to_host(aC, len(C0)), want_add[0], f"ge_add {name} L={L}")
        same(c.array_to_host(buf, m * n), want_add[1], f"ge_add buffer {name} L={L}")
        mp.check(c.L.mpres_ge_acc(c.h, m, n, C.byref(aal), C.byref(aA), lda, C.byref(abe), C.byref(aB), ldb,
                                  C.byref(buf), None))
        same(c.array_to_host(aB, len(B0)), want_acc[0], f"ge_acc {name} L={L}")
        aD = A(D0)
        for (side, inc), want in zip(diag_cases, want_diag):
            aA = A(A0)
            mp.check(c.L.mpres_ge_diag_scale(c.h, side, m, n, C.byref(aD), inc, C.byref(aA), lda, None))
            same(c.array_to_host(aA, len(A0)), want, f"ge_diag_scale {name} L={L} side {side}")
        aA = A(A0)
        mp.check(c.L.mpres_ge_lr_scale(c.h, m, n, C.byref(aD), -2, C.byref(aD), 1, C.byref(aA), lda, None))
        same(c.array_to_host(aA, len(A0)), want_lr, f"ge_lr_scale {name} L={L}")
        aA = A(A0)
        for (nm, g, b), (wr, wsums) in zip(norm_cases, want_norm):
            outs = n if nm == mp.ONE_NORM else m
            ar, sb = c.array(1), c.array(outs)
            mp.check(c.L.mpres_ge_norm(c.h, nm, m, n, C.byref(aA), lda, C.byref(ar), C.byref(sb), g, b, None, 0,
                                       None))
            same(c.array_to_host(sb, outs), wsums, f"ge_norm sums {name} L={L} norm {nm}")
            same(c.array_to_host(ar, 1), wr, f"ge_norm {name} L={L} norm {nm}")
        for (ix, iy), (wa, wb1, wb2) in zip(ger_cases, want_ger):
            aA, ax, ay, b1, b2 = A(A0), A(x0), A(y0), c.array(n), c.array(m * n)
            mp.check(c.L.mpres_ger(c.h, m, n, C.byref(aal), C.byref(ax), ix, C.byref(ay), iy, C.byref(aA), lda,
                                   C.byref(b1), C.byref(b2), None))
            same(c.array_to_host(aA, len(A0)), wa, f"ger {name} L={L} inc {ix},{iy}")
            same(c.array_to_host(b1, n), wb1, f"ger buffer1 {name} L={L}")
            same(c.array_to_host(b2, m * n), wb2, f"ger buffer2 {name} L={L}")


@pytest.mark.parametrize("name", SETS)
def test_v2_routines(name):
    c, o = ctx(name), orc(name)
    n, lda = 45, 47
    A0, x0, y0 = vals(o, 51, lda * n), vals(o, 53, 3 * n), vals(o, 55, 2 * n)
    al, be = vals(o, 57, 2)[1:2], scalar(o, 1.25)
    want_scal = o.scal_v2(al, x0)
    want_prod = o.mul(x0[:2 * n], y0, D)
    m = 30
    want_ger = o.ger_v2(m, n, al, x0, -3, y0, 2, A0, lda)
    cases = [True, False]
    want_syr = [o.syr_v2(u, n, al, x0, 3, A0, lda) for u in cases]
    want_syr2 = [o.syr2_v2(u, n, al, x0, -2, y0, 1, A0, lda) for u in cases]
    want_symv = [o.symv_v2(u, n, al, A0, lda, x0, 2, be, y0, -2) for u in cases]
    k, ldc = 17, 46
    C0 = vals(o, 59, ldc * n)
    syrk_cases = [(u, t) for u in cases for t in (mp.NO_TRANS, mp.TRANS, mp.CONJ_TRANS)]
    want_syrk = [o.syrk_v2(u, t == mp.TRANS, n, k, al, A0, lda, be, C0, ldc) for u, t in syrk_cases]
    for L in lanes_of(c):
        c.set_lanes(mp.K_ELEM, L)
        note("v2 routines", name, L)
        dA, dx, dy, dal, dbe = c.aos(A0), c.aos(x0), c.aos(y0), c.aos(al), c.aos(be)
        dr = c.aos(count=len(x0))
        mp.check(c.L.mpres_scal_v2(c.h, len(x0), dal.ptr, dx.ptr, dr.ptr, None))
        same(c.aos_to_host(dr, len(x0)), want_scal, f"scal_v2 {name} L={L}")
        mp.check(c.L.mpres_prod_v2(c.h, 2 * n, dx.ptr, dy.ptr, dr.ptr, None))
        same(c.aos_to_host(dr, 2 * n), want_prod, f"prod_v2 {name} L={L}")
        mp.check(c.L.mpres_ger_v2(c.h, m, n, dal.ptr, dx.ptr, -3, dy.ptr, 2, dA.ptr, lda, None))
        same(c.aos_to_host(dA, len(A0)), want_ger, f"ger_v2 {name} L={L}")
        for u, w1, w2, w3 in zip(cases, want_syr, want_syr2, want_symv):
            up = mp.UPPER if u else mp.LOWER
            dA = c.aos(A0)
            mp.check(c.L.mpres_syr_v2(c.h, up, n, dal.ptr, dx.ptr, 3, dA.ptr, lda, None))
            same(c.aos_to_host(dA, len(A0)), w1, f"syr_v2 {name} L={L} uplo {up}")
            dA = c.aos(A0)
            mp.check(c.L.mpres_syr2_v2(c.h, up, n, dal.ptr, dx.ptr, -2, dy.ptr, 1, dA.ptr, lda, None))
            same(c.aos_to_host(dA, len(A0)), w2, f"syr2_v2 {name} L={L} uplo {up}")
            dA, dyy = c.aos(A0), c.aos(y0)
            mp.check(c.L.mpres_symv_v2(c.h, up, n, dal.ptr, dA.ptr, lda, dx.ptr, 2, dbe.ptr, dyy.ptr, -2, None))
            same(c.aos_to_host(dyy, len(y0)), w3, f"symv_v2 {name} L={L} uplo {up}")
        dA = c.aos(A0)
        for (u, t), want in zip(syrk_cases, want_syrk):
            dC = c.aos(C0)
            mp.check(c.L.mpres_syrk_v2(c.h, mp.UPPER if u else mp.LOWER, t, n, k, dal.ptr, dA.ptr, lda, dbe.ptr,
                                       dC.ptr, ldc, None))
            same(c.aos_to_host(dC, len(C0)), want, f"syrk_v2 {name} L={L} upper {u} trans {t}")


def test_argument_errors():
    c = ctx("8_2double")
    a = c.array(4)
    bad = [
        lambda: c.L.mpres_ge_add(c.h, 4, 2, C.byref(a), C.byref(a), 3, C.byref(a), C.byref(a), 4, C.byref(a), 4,
                                 C.byref(a), None),
        lambda: c.L.mpres_ge_diag_scale(c.h, mp.LEFT, 2, 2, C.byref(a), 0, C.byref(a), 2, None),
        lambda: c.L.mpres_ge_diag_scale(c.h, 7, 2, 2, C.byref(a), 1, C.byref(a), 2, None),
        lambda: c.L.mpres_norm(c.h, 3, 2, C.byref(a), 1, C.byref(a), 1, 1, None, 0, None),
        lambda: c.L.mpres_ger(c.h, 2, 2, C.byref(a), C.byref(a), 0, C.byref(a), 1, C.byref(a), 2, C.byref(a),
                              C.byref(a), None),
        lambda: c.L.mpres_syr_v2(c.h, 5, 2, None, None, 1, None, 2, None),
    ]
    for f in bad:
        assert f() == -1  # MPRES_ERR_ARG
    # quick returns (the reference's): nothing launched, success
    assert c.L.mpres_scal(c.h, 0, C.byref(a), C.byref(a), 1, None) == 0
    assert c.L.mpres_asum(c.h, 4, C.byref(a), 0, C.byref(a), 1, 1, None, 0, None) == 0
