"""GPU parity: every entry point of the HIP library against the oracle's
restatement of the reference CUDA algorithm (policy D), bit for bit in every
field (residues, sign, exponent, both interval bounds), on seeded inputs at
sizes the oracle finishes in seconds, for every compiled lanes-per-number
variant."""
import ctypes as C
import math
import importlib.util
import os
import sys

import numpy as np
import pytest

import exact
from fractions import Fraction
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
_ctx = {}
_orc = {}


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
    got = np.ascontiguousarray(got)
    want = np.ascontiguousarray(want)
    assert got.dtype == want.dtype and got.shape == want.shape
    if got.tobytes() == want.tobytes():
        return
    bad = [i for i in range(len(got)) if got[i].tobytes() != want[i].tobytes()]
    i = bad[0]
    raise AssertionError(f"{what}: {len(bad)} of {len(got)} records differ; first {i}:\n gpu {got[i]}\n ref {want[i]}")


SETS = ["8_2double", "16_4double", "32_8double", "8_32bit"]


def inputs(o, seed, count, elo=-40, ehi=40):
    x = gen.uexp(seed, count, elo, ehi)
    x[:6] = [0.0, -0.0, 1.0, -1.0, 0.75, -3.0]
    return x


@pytest.mark.parametrize("name", SETS)
def test_set_d(name):
    c, o = ctx(name), orc(name)
    x = np.concatenate([inputs(o, 1, 3000, -300, 300), [2.0**-1000, -2.0**900, 5e-324]])
    for L in lanes_of(c):
        note("set_d", name, L)
        c.set_lanes(mp.K_ELEM, L)
        got = c.aos_to_host(c.set_d(x), len(x))
        same(got, o.set_d(x, D), f"set_d {name} L={L}")


def _operands(o, seed, cnt):
    a = o.set_d(inputs(o, seed, cnt), D)
    b = o.set_d(inputs(o, seed + 1, cnt), D)
    # products carry 106-bit significands: at 8 moduli their sums round
    pa = o.mul(a, b, D)
    pb = o.mul(b[::-1].copy(), a, D)
    # cancellations: exact (x + -x) and near-exact, to reach the sign estimate (MRC) path
    neg = pa.copy()
    neg["sign"] ^= 1
    near = o.mul(pa, o.set_d(np.full(cnt, 1.0 + 2.0**-40), D), D)
    near["sign"] ^= 1
    X = np.concatenate([a, pa, pa, pa, pa])
    Y = np.concatenate([b, pb, neg, near, pb[::-1]])
    return X, Y


@pytest.mark.parametrize("name", SETS)
def test_elementwise(name):
    c, o = ctx(name), orc(name)
    X, Y = _operands(o, 11, 600)
    yd = inputs(o, 13, len(X), -30, 30)
    dX, dY, dYd = c.aos(X), c.aos(Y), mp.DevBuf(host=yd)
    want = {mp.OP_ADD: o.add(X, Y, D), mp.OP_SUB: o.sub(X, Y, D), mp.OP_MUL: o.mul(X, Y, D),
            mp.OP_MUL_D: o.mul_d(X, yd, D), mp.OP_VADD: o.vadd(X, Y)}
    for L in lanes_of(c):
        c.set_lanes(mp.K_ELEM, L)
        for op, w in want.items():
            note("elementwise", name, L, op)
            r = c.aos(count=len(X))
            mp.check(c.L.mpres_elementwise(c.h, op, len(X), dX.ptr, dYd.ptr if op == mp.OP_MUL_D else dY.ptr, r.ptr,
                                           None))
            same(c.aos_to_host(r, len(X)), w, f"op {op} {name} L={L}")


def _cmp_operands(o, seed, cnt):
    """pairs reaching every branch of mp_cmp: separated evaluations, equal numbers (also
    with different representations of one value), near-equal magnitudes (overlapping
    evaluations, mixed-radix decision), signed zeros, opposite signs, an operand dropped
    against a much larger one"""
    X, Y = _operands(o, seed, cnt)
    a = o.set_d(inputs(o, seed + 7, cnt, -300, 300), D)
    neg = a.copy()
    neg["sign"] ^= 1
    z = o.set_d(np.array([0.0, -0.0] * (cnt // 2)), D)
    tiny = o.mul(a, o.set_d(np.full(cnt, 2.0**-600), D), D)
    # the same values scaled by a power of two through the residues: equal, other exponent
    two = o.mul(o.mul(a, o.set_d(np.full(cnt, 2.0**40), D), D), o.set_d(np.full(cnt, 2.0**-40), D), D)
    up = o.add(a, o.mul(a, o.set_d(np.full(cnt, 2.0**-90), D), D), D)
    Xs = np.concatenate([X, a, a, a, z, z[::-1], a, tiny, two, up, a])
    Ys = np.concatenate([Y, a, neg, z, z, a, tiny, a, a, a, up])
    return Xs, Ys


@pytest.mark.parametrize("name", SETS)
def test_cmp(name):
    """mpres_cmp (mp_cmp, mp_cmp_abs) and mpres_rns_cmp against the oracle, every lanes variant"""
    c, o = ctx(name), orc(name)
    X, Y = _cmp_operands(o, 17, 400)
    dX, dY = c.aos(X), c.aos(Y)
    want = {0: o.cmp(X, Y, False, D), 1: o.cmp(X, Y, True, D)}
    assert set(np.unique(want[0])) == {-1, 0, 1}
    # bare residues: random integers below M, equal pairs, neighbours (X, X + 1), small X
    mm = [int(v) for v in o.moduli()]
    M = exact.moduli_product(mm)
    rng = np.random.default_rng(19)
    ints = [int(rng.integers(0, 2**62)) * int(rng.integers(0, 2**62)) % M for _ in range(300)]
    ints += [int(rng.integers(0, 2**40)) for _ in range(100)]
    pairs = [(a_, b_) for a_, b_ in zip(ints[:200], ints[200:])]
    pairs += [(v, v) for v in ints[:50]] + [(v, (v + 1) % M) for v in ints[50:100]] + [((v + 1) % M, v) for v in ints[100:150]]
    pairs += [(M - 1, M - 2), (0, 0), (0, 1), (M - 1, 0)]
    dx = np.array([exact.residues(a_, mm) for a_, _ in pairs], np.int32)
    dy = np.array([exact.residues(b_, mm) for _, b_ in pairs], np.int32)
    want_r = o.rns_cmp(dx, dy, D)
    assert list(want_r) == [(a_ > b_) - (a_ < b_) for a_, b_ in pairs]
    ddx, ddy = mp.DevBuf(host=dx), mp.DevBuf(host=dy)
    for L in lanes_of(c):
        c.set_lanes(mp.K_ELEM, L)
        for mode, w in want.items():
            note("cmp", name, mode, L)
            r = mp.DevBuf(nbytes=4 * len(X))
            mp.check(c.L.mpres_cmp(c.h, mode, len(X), dX.ptr, dY.ptr, r.ptr, None))
            got = r.to_numpy(np.int32, len(X))
            bad = np.nonzero(got != w)[0]
            assert len(bad) == 0, f"cmp {name} mode {mode} L={L}: {len(bad)} differ, first {bad[:5]}"
        note("rns_cmp", name, L)
        r = mp.DevBuf(nbytes=4 * len(pairs))
        mp.check(c.L.mpres_rns_cmp(c.h, len(pairs), ddx.ptr, ddy.ptr, r.ptr, None))
        got = r.to_numpy(np.int32, len(pairs))
        assert (got == want_r).all(), f"rns_cmp {name} L={L}: {np.nonzero(got != want_r)[0][:5]}"


@pytest.mark.parametrize("name,n,g3,b3", [("8_2double", 5000, 64, 64), ("8_2double", 3001, 256, 64),
                                          ("8_2double", 700, 7, 33), ("16_4double", 2000, 64, 64),
                                          ("32_8double", 1500, 32, 64)])
def test_dot(name, n, g3, b3):
    c, o = ctx(name), orc(name)
    x = o.set_d(gen.uexp(21, n, -20, 20), D)
    y = o.set_d(gen.uexp(22, n, -20, 20), D)
    ax, ay = c.array_from_host(x), c.array_from_host(y)
    r = c.array(1)
    dx, dy, dr = c.aos(x), c.aos(y), c.aos(count=1)
    want = o.dot(x, y, g3, b3, D)
    want_s = o.dot(x, y, g3, b3, D, incx=-1, incy=2, n=n // 2)
    for L in lanes_of(c):
        c.set_lanes(mp.K_DOT, L)
        note("dot", name, n, g3, b3, L)
        mp.check(c.L.mpres_dot(c.h, n, C.byref(ax), 1, C.byref(ay), 1, C.byref(r), g3, b3, None, 0, None))
        same(c.array_to_host(r, 1), want, f"dot {name} L={L}")
        mp.check(c.L.mpres_dot(c.h, n // 2, C.byref(ax), -1, C.byref(ay), 2, C.byref(r), g3, b3, None, 0, None))
        same(c.array_to_host(r, 1), want_s, f"dot strided {name} L={L}")
        mp.check(c.L.mpres_dot_v2(c.h, n, dx.ptr, dy.ptr, dr.ptr, g3, b3, None, 0, None))
        same(c.aos_to_host(dr, 1), want, f"dot_v2 {name} L={L}")
    # the result is an exactly rounded-down sum of exactly rounded products: compare with the exact value
    m = [int(v) for v in o.moduli()]
    ex = sum(exact.mp_value(p, m) for p in o.mul(x, y, D))
    got = exact.mp_value(want[0], m)
    if ex != 0:
        assert abs(got - ex) / abs(ex) < 2.0 ** (-c.precision + 12)


@pytest.mark.parametrize("name", ["8_2double", "32_8double"])
def test_lanes_table(name):
    """mpres_set_lanes_table: calls pick their lanes by size (first bound at or above the size,
    the last entry beyond), results stay those of the oracle, bad tables are refused and an
    empty table restores the per-operation choice"""
    c, o = ctx(name), orc(name)
    lanes = lanes_of(c)
    small, large = lanes[0], lanes[-1]
    default = c.L.mpres_get_lanes(c.h, mp.K_DOT)
    try:
        c.set_lanes_table(mp.K_DOT, [(1000, small), (100000, large)])
        assert (c.lanes_for(mp.K_DOT, 1), c.lanes_for(mp.K_DOT, 1000), c.lanes_for(mp.K_DOT, 1001),
                c.lanes_for(mp.K_DOT, 10 ** 12)) == (small, small, large, large)
        for n, seed in ((700, 61), (5000, 62)):
            note("lanes table dot", name, n)
            x, y = o.set_d(gen.uexp(seed, n, -20, 20), D), o.set_d(gen.uexp(seed + 1, n, -20, 20), D)
            ax, ay, r = c.array_from_host(x), c.array_from_host(y), c.array(1)
            mp.check(c.L.mpres_dot(c.h, n, C.byref(ax), 1, C.byref(ay), 1, C.byref(r), 16, 32, None, 0, None))
            same(c.array_to_host(r, 1), o.dot(x, y, 16, 32, D), f"dot {name} n={n} with a lanes table")
        B = C.c_longlong * 2
        Lx = C.c_int * 2
        assert c.L.mpres_set_lanes_table(c.h, mp.K_DOT, 2, B(10, 10), Lx(small, small)) == mp.MPRES_ERR_ARG
        assert c.L.mpres_set_lanes_table(c.h, mp.K_DOT, 1, B(10, 0), Lx(3, 0)) == mp.MPRES_ERR_UNSUPPORTED
        assert c.L.mpres_set_lanes_table(c.h, mp.K_DOT, 9, B(1, 2), Lx(small, small)) == mp.MPRES_ERR_ARG
        assert c.L.mpres_set_lanes_table(c.h, 6, 1, B(1, 2), Lx(small, small)) == mp.MPRES_ERR_ARG
        c.set_lanes_table(mp.K_DOT, [])
        assert c.lanes_for(mp.K_DOT, 1000) == c.lanes_for(mp.K_DOT, 10 ** 9) == default
        c.set_lanes(mp.K_DOT, small)  # a fixed choice replaces any table
        assert c.lanes_for(mp.K_DOT, 10 ** 9) == small
    finally:
        c.set_lanes(mp.K_DOT, 0)  # the built-in choice again


def test_builtin_lanes_tables():
    """the measured lanes tables installed at context creation (106-bit gemv of 64^2 on 8
    lanes and of 4096^2 on one, gemm of 512^3 on one), set_lanes overriding them and 0
    restoring them"""
    c8 = ctx("8_2double")
    try:
        c8.set_lanes(mp.K_GEMV, 0)
        assert (c8.lanes_for(mp.K_GEMV, 64 * 64), c8.lanes_for(mp.K_GEMV, 4096 * 4096)) == (8, 1)
        c8.set_lanes(mp.K_GEMV, 2)
        assert c8.lanes_for(mp.K_GEMV, 4096 * 4096) == 2
        c8.set_lanes(mp.K_GEMV, 0)
        assert c8.lanes_for(mp.K_GEMV, 4096 * 4096) == 1
        c8.set_lanes(mp.K_GEMM, 0)
        assert c8.lanes_for(mp.K_GEMM, 512 ** 3) == 1
    finally:
        c8.set_lanes(mp.K_GEMV, 0)
        c8.set_lanes(mp.K_GEMM, 0)


def _get_d_exact(rec, m):
    """(-1)^sign X 2^exp as the reference's MPFR conversion gives it: X rounded to 53 bits to
    nearest (ties to even), then to a double (C ldexp: one more rounding when subnormal)"""
    X = exact.crt(rec["digits"], m)
    s, e = int(rec["sign"]), int(rec["exp"])
    if X == 0:
        return -0.0 if s else 0.0
    b = X.bit_length()
    if b > 53:
        q, r = divmod(X, 1 << (b - 53))
        h = 1 << (b - 54)
        if r > h or (r == h and q & 1):
            q += 1
        X, e = q, e + b - 53
    try:
        v = math.ldexp(float(X), e)
    except OverflowError:
        v = math.inf
    return -v if s else v


@pytest.mark.parametrize("name", SETS)
def test_get_d_and_integer(name):
    """mpres_get_d / mpres_array_get_d / mpres_get_integer (the reference's mp_get_d and
    mp_get_mpfr on the device): round trips of doubles, random residues scaled into the
    normal, subnormal and overflow ranges, and products of full-length numbers"""
    c, o = ctx(name), orc(name)
    m = [int(v) for v in o.moduli()]
    M = exact.moduli_product(m)
    rng = np.random.default_rng(81)
    # 1. round trips: every double survives set_d -> get_d (zeros come back as +0, as set_d stores them)
    xd = np.concatenate([inputs(o, 82, 2000, -1000, 1000), [2.0**-1074, -2.0**-1060, 2.0**1023 * 1.5, -0.0,
                                                             np.nextafter(2.0**-1022, 0), 1.0 - 2.0**-53]])
    recs = o.set_d(xd, D)
    # 2. random X in [0, M) with exponents that put |X 2^e| in the normal, subnormal and overflow ranges
    cnt = 1500
    rnd = np.zeros(cnt, o.dtype)
    for i in range(cnt):
        X = int.from_bytes(rng.bytes((M.bit_length() + 7) // 8), "little") % M
        if i % 7 == 0:
            X >>= int(rng.integers(0, M.bit_length()))  # short integers too
        b = max(X.bit_length(), 1)
        kind = i % 3
        e = int(rng.integers(-900, 900)) - b if kind == 0 else (-1074 - b + int(rng.integers(-5, 60)) if kind == 1
                                                                 else 1024 - b + int(rng.integers(-2, 3)))
        rnd[i]["digits"] = [X % mi for mi in m]
        rnd[i]["sign"] = int(rng.integers(0, 2))
        rnd[i]["exp"] = e
    # ties: X = (2^53 + 1) 2^k + 2^(k-1) patterns and exact halves
    for j, X in enumerate([(2**53 + 1) * 2**10 + 2**9, 2**53 * 2**10 + 2**9, (2**53 + 3) * 2**4 + 2**3 + 1]):
        if X < M:
            rnd[j]["digits"] = [X % mi for mi in m]
            rnd[j]["exp"] = -20
    # 3. products of full-length numbers (op 2: mp_mul)
    a, b = o.set_d(inputs(o, 83, 500), D), o.set_d(inputs(o, 84, 500), D)
    prods = o.mul(a, b, D)
    allr = np.concatenate([recs, rnd, prods])
    n = len(allr)
    want = np.array([_get_d_exact(r, m) for r in allr])
    dx = c.aos(allr)
    out = mp.DevBuf(nbytes=8 * n)
    note("get_d", name, n)
    mp.check(c.L.mpres_get_d(c.h, n, dx.ptr, out.ptr, None))
    got = out.to_numpy(np.float64, n)
    assert got.tobytes() == want.tobytes(), f"{int((got.view(np.int64) != want.view(np.int64)).sum())} differ"
    nz = np.abs(xd) >= 2.0**-1022  # normal doubles round-trip exactly (set_d stores no subnormals)
    assert np.array_equal(got[:len(xd)][nz], xd[nz])
    # the same through the mp_array_t layout
    arr = c.array_from_host(allr)
    out2 = mp.DevBuf(nbytes=8 * n)
    mp.check(c.L.mpres_array_get_d(c.h, n, C.byref(arr), out2.ptr, None))
    assert out2.to_numpy(np.float64, n).tobytes() == want.tobytes()
    # exact integers
    lw = c.L.mpres_export_limbs(c.h)
    lb, sg, ex = mp.DevBuf(nbytes=4 * lw * n), mp.DevBuf(nbytes=4 * n), mp.DevBuf(nbytes=4 * n)
    mp.check(c.L.mpres_get_integer(c.h, n, dx.ptr, lb.ptr, sg.ptr, ex.ptr, None))
    limbs = lb.to_numpy(np.uint32, lw * n).reshape(n, lw)
    assert np.array_equal(sg.to_numpy(np.int32, n), allr["sign"])
    assert np.array_equal(ex.to_numpy(np.int32, n), allr["exp"])
    for i in range(n):
        X = int.from_bytes(limbs[i].tobytes(), "little")
        assert X == exact.crt(allr[i]["digits"], m), f"record {i}"


@pytest.mark.parametrize("name", ["8_2double", "16_4double", "32_8double"])
def test_div_sqrt(name):
    """mpres_div / mpres_sqrt / mpres_dsqrt (the reference's host MPFR operations) against the
    oracle's MPFR restatement, bit for bit: full-length significands, zeros of both signs,
    equal operands, powers of two"""
    c, o = ctx(name), orc(name)
    a = o.set_d(inputs(o, 91, 300), D)
    b = o.set_d(inputs(o, 92, 300), D)
    x = np.concatenate([o.mul(a, b, D), o.add(a, b, D), a, o.set_d([0.0, -0.0, 4.0, 2.0**-30], D)])
    y = np.concatenate([o.add(b, a, D), o.mul(b, b, D), b[::-1], o.set_d([3.0, -3.0, 4.0, 1.0], D)])
    y[y["eval"][:, 1]["frac"] == 0] = o.set_d([7.0], D)[0]  # no zero divisors
    n = len(x)
    note("div/sqrt", name, n)
    dx, dy, dr = c.aos(x), c.aos(y), c.aos(count=n)
    mp.check(c.L.mpres_div(c.h, n, dx.ptr, dy.ptr, dr.ptr, None))
    same(c.aos_to_host(dr, n), o.div(x, y), f"div {name}")
    xs = np.array(x, copy=True)
    xs["sign"] = 0
    ds, dd = c.aos(xs), mp.DevBuf(nbytes=8 * n)
    mp.check(c.L.mpres_sqrt(c.h, n, ds.ptr, dr.ptr, None))
    same(c.aos_to_host(dr, n), o.sqrt(xs), f"sqrt {name}")
    mp.check(c.L.mpres_dsqrt(c.h, n, ds.ptr, dd.ptr, None))
    assert dd.to_numpy(np.float64, n).tobytes() == o.dsqrt(xs).tobytes()


@pytest.mark.parametrize("name", ["8_2double", "32_8double"])
def test_solver_helpers(name):
    """v2 maxpy, diff, prod_d and norm2 (every order of the two reduction kernels used)"""
    c, o = ctx(name), orc(name)
    n = 3000
    x, y = o.set_d(inputs(o, 93, n), D), o.set_d(inputs(o, 94, n), D)
    alpha = o.set_d([-0.375], D)
    dvals = gen.uexp(95, n, -10, 10)
    dx, dy, da, dr = c.aos(x), c.aos(y), c.aos(alpha), c.aos(count=n)
    note("solver helpers", name)
    mp.check(c.L.mpres_maxpy_v2(c.h, n, da.ptr, dx.ptr, dy.ptr, dr.ptr, None))
    same(c.aos_to_host(dr, n), o.maxpy_v2(alpha, x, y, D), f"maxpy {name}")
    mp.check(c.L.mpres_diff_v2(c.h, n, dx.ptr, dy.ptr, dr.ptr, None))
    same(c.aos_to_host(dr, n), o.sub(x, y, D), f"diff {name}")
    dd = mp.DevBuf(host=dvals)
    mp.check(c.L.mpres_prod_d_v2(c.h, n, dx.ptr, dd.ptr, dr.ptr, None))
    same(c.aos_to_host(dr, n), o.mul_d(x, dvals, D), f"prod_d {name}")
    for blocks, threads in ((256, 32), (7, 64), (3, 5)):
        r = C.c_double()
        mp.check(c.L.mpres_norm2_v2(c.h, n, dx.ptr, blocks, threads, C.byref(r), None, 0, None))
        assert r.value == o.norm2(x, blocks, threads, D), f"norm2 {name} {blocks}x{threads}"


def _cg_case(name, n, irp, ja, as_, b, tol, maxit, precond, threads, blocks_reduce):
    c, o = ctx(name), orc(name)
    x0 = o.set_d(np.zeros(n), D)
    xw, rw = o.cg_csr(n, irp, ja, as_, b, tol, maxit, x0, threads=threads, blocks_reduce=blocks_reduce,
                      precond=precond)
    dirp, dja, das = mp.DevBuf(host=irp), mp.DevBuf(host=ja), mp.DevBuf(host=as_)
    csr = mp.Csr(das.ptr.value, dja.ptr.value, dirp.ptr.value)
    db, dx = c.aos(b), c.aos(x0)
    dm = mp.DevBuf(host=precond) if precond is not None else None
    res = (C.c_double * (maxit + 1))()
    it = C.c_int()
    mp.check(c.L.mpres_cg_csr(c.h, n, C.byref(csr), db.ptr, tol, maxit, dx.ptr, dm.ptr if dm else None, res,
                              C.byref(it), threads, blocks_reduce, None, 0, None))
    assert it.value == len(rw) - 1, f"{it.value} iterations, oracle {len(rw) - 1}"
    assert list(res[:it.value + 1]) == rw
    same(c.aos_to_host(dx, n), xw, f"cg x {name}")
    return it.value


@pytest.mark.parametrize("name", ["8_2double", "16_4double", "32_8double"])
@pytest.mark.parametrize("precond", [False, True])
def test_cg_lf10(name, precond):
    """mp_cg_csr / mp_pcg_csr (diagonal) on the reference's LF10 matrix: every iterate's
    relative residual and the solution, bit for bit against the oracle's composition"""
    n, _, irp, ja, as_ = gen.read_mtx(os.path.join(ROOT, "tests", "golden", "LF10.mtx"))
    o = orc(name)
    b = o.spmv(n, irp, ja, as_, o.set_d(gen.uexp(96, n, -2, 2), D), 1, D)
    diag = np.array([as_[irp[i] + list(ja[irp[i]:irp[i + 1]]).index(i)] for i in range(n)])
    note("cg lf10", name, precond)
    its = _cg_case(name, n, irp, ja, as_, b, {"8_2double": 1e-25, "16_4double": 1e-50, "32_8double": 1e-90}[name],
                   300, 1.0 / diag if precond else None, 32, 256)
    assert its > 5


def test_cg_laplacian():
    """CG on the 1D Laplacian of size 5000 (several blocks in every reduction), 30 iterations"""
    n = 5000
    irp = np.zeros(n + 1, np.int32)
    ja, as_ = [], []
    for i in range(n):
        for j, v in ((i - 1, -1.0), (i, 2.0), (i + 1, -1.0)):
            if 0 <= j < n:
                ja.append(j)
                as_.append(v)
        irp[i + 1] = len(ja)
    ja, as_ = np.array(ja, np.int32), np.array(as_, np.float64)
    o = orc("16_4double")
    b = o.set_d(gen.uexp(97, n, -3, 3), D)
    note("cg laplacian", n)
    _cg_case("16_4double", n, irp, ja, as_, b, 1e-40, 30, None, 64, 128)


def _mtx_all(path):
    """the library's Matrix Market reader: CSR, ELL, JAD and DIA host arrays of one file"""
    L = mp.lib()
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    v = [C.c_int() for _ in range(6)]
    mp.check(L.mpres_mtx_info(path.encode(), *[C.byref(q) for q in v]))
    m, n, lines, maxnzr, nzmd, sym = (q.value for q in v)
    nnz = L.mpres_mtx_nnz(lines, nzmd, sym)
    csr = (np.zeros(nnz), np.zeros(nnz, np.int32), np.zeros(m + 1, np.int32))
    mp.check(L.mpres_mtx_csr(path.encode(), m, nnz, lines, sym, *[P(a) for a in csr]))
    ell = (np.zeros(m * maxnzr), np.zeros(m * maxnzr, np.int32))
    mp.check(L.mpres_mtx_ell(path.encode(), m, maxnzr, lines, sym, *[P(a) for a in ell]))
    jad = (np.zeros(nnz), np.zeros(nnz, np.int32), np.zeros(maxnzr + 1, np.int32), np.zeros(m, np.int32))
    mp.check(L.mpres_mtx_jad(path.encode(), m, maxnzr, nnz, lines, sym, *[P(a) for a in jad]))
    nd = C.c_int()
    mp.check(L.mpres_mtx_ndiag(path.encode(), lines, sym, C.byref(nd)))
    dia = (np.zeros(m * nd.value), np.zeros(nd.value, np.int32))
    mp.check(L.mpres_mtx_dia(path.encode(), m, lines, sym, nd.value, *[P(a) for a in dia]))
    return dict(m=m, n=n, maxnzr=maxnzr, ndiag=nd.value, csr=csr, ell=ell, jad=jad, dia=dia)


class _Dev:
    """device copies of a format's host arrays and the struct that points at them"""

    def __init__(self, kind, arrays):
        self.bufs = [mp.DevBuf(host=a) for a in arrays]
        ptrs = [b.ptr.value for b in self.bufs]
        self.s = {"csr": mp.Csr, "ell": mp.Ell, "jad": mp.Jad, "dia": mp.Dia}[kind](*ptrs)


def _spmv_all(c, o, S, x, tprs=(2, 8, 32)):
    """every format's GPU SpMV against the oracle on the same arrays; returns the device structs"""
    m, w, nd = S["m"], S["maxnzr"], S["ndiag"]
    dx, dy = c.aos(x), c.aos(count=m)
    dev = {k: _Dev(k, S[k]) for k in ("csr", "ell", "jad", "dia")}
    mp.check(c.L.mpres_spmv_ell(c.h, m, w, C.byref(dev["ell"].s), dx.ptr, dy.ptr, None))
    same(c.aos_to_host(dy, m), o.spmv_ell(m, w, *S["ell"], x, D), "ell")
    mp.check(c.L.mpres_spmv_jad(c.h, m, w, C.byref(dev["jad"].s), dx.ptr, dy.ptr, None))
    same(c.aos_to_host(dy, m), o.spmv_jad(m, w, *S["jad"], x, 1, D), "jad")
    for tpr in tprs:
        mp.check(c.L.mpres_spmv_jadv(c.h, m, w, C.byref(dev["jad"].s), tpr, dx.ptr, dy.ptr, None))
        same(c.aos_to_host(dy, m), o.spmv_jad(m, w, *S["jad"], x, tpr, D), f"jadv tpr={tpr}")
    if nd:
        mp.check(c.L.mpres_spmv_dia(c.h, m, S["n"], nd, C.byref(dev["dia"].s), dx.ptr, dy.ptr, None))
        same(c.aos_to_host(dy, m), o.spmv_dia(m, S["n"], nd, *S["dia"], x, D), "dia")
    return dev


@pytest.mark.parametrize("name", ["8_2double", "16_4double", "32_8double"])
def test_spmv_formats_lf10(name):
    """ELL / JAD / JADV / DIA on the reference's LF10 file through the library's reader"""
    c, o = ctx(name), orc(name)
    S = _mtx_all(os.path.join(ROOT, "tests", "golden", "LF10.mtx"))
    note("spmv formats lf10", name)
    _spmv_all(c, o, S, o.set_d(gen.uexp(101, S["m"], -20, 20), D))


@pytest.mark.parametrize("name", ["8_2double", "16_4double", "32_8double"])
def test_spmv_formats_large(name):
    """20000 power-law rows in ELL and JAD (rows with exactly cancelling pairs take the fixup
    path), and a banded 20000-row matrix in DIA"""
    c, o = ctx(name), orc(name)
    m = 20000
    irp, ja, as_ = gen.powerlaw_csr(102, m, kmax=64)
    xd = gen.uexp(103, m, -30, 30)
    for i in range(0, m, 97):  # a pair of equal x entries with opposite coefficients in the row
        if irp[i + 1] - irp[i] >= 2:
            a, b = ja[irp[i]], ja[irp[i] + 1]
            xd[b] = xd[a]
            as_[irp[i] + 1] = -as_[irp[i]]
    x = o.set_d(xd, D)
    lens = np.diff(irp)
    w = int(lens.max())
    eas, eja = np.zeros(m * w), np.full(m * w, -1, np.int32)
    for i in range(m):
        eas[np.arange(lens[i]) * m + i] = as_[irp[i]:irp[i + 1]]
        eja[np.arange(lens[i]) * m + i] = ja[irp[i]:irp[i + 1]]
    perm = np.argsort(-lens, kind="stable").astype(np.int32)
    jcp = np.concatenate([[0], np.cumsum([(lens > j).sum() for j in range(w)])]).astype(np.int32)
    jas = np.concatenate([as_[irp[perm[:jcp[j + 1] - jcp[j]]] + j] for j in range(w)])
    jja = np.concatenate([ja[irp[perm[:jcp[j + 1] - jcp[j]]] + j] for j in range(w)]).astype(np.int32)
    offs = np.arange(-8, 9, dtype=np.int32)
    das = gen.uexp(104, m * len(offs), 0, 0) * (np.arange(m * len(offs)) % 5 != 0)  # some zeros
    S = dict(m=m, n=m, maxnzr=w, ndiag=len(offs), csr=(as_, ja, irp), ell=(eas, eja), jad=(jas, jja, jcp, perm),
             dia=(das, offs))
    note("spmv formats large", name, m)
    _spmv_all(c, o, S, x)


@pytest.mark.parametrize("name", ["8_2double", "32_8double"])
@pytest.mark.parametrize("fmt", ["ell", "jad", "dia"])
def test_cg_formats_lf10(name, fmt):
    """mp_cg_ell / mp_cg_jad / mp_cg_dia (and the diagonal PCG) on LF10 through the reader,
    every iterate bit-exact against the oracle's composition over the same format"""
    c, o = ctx(name), orc(name)
    S = _mtx_all(os.path.join(ROOT, "tests", "golden", "LF10.mtx"))
    n = S["m"]
    irp, ja, as_ = S["csr"][2], S["csr"][1], S["csr"][0]
    b = o.spmv(n, irp, ja, as_, o.set_d(gen.uexp(105, n, -2, 2), D), 1, D)
    diag = np.array([as_[irp[i] + list(ja[irp[i]:irp[i + 1]]).index(i)] for i in range(n)])
    spmv = {"ell": lambda v: o.spmv_ell(n, S["maxnzr"], *S["ell"], v, D),
            "jad": lambda v: o.spmv_jad(n, S["maxnzr"], *S["jad"], v, 1, D),
            "dia": lambda v: o.spmv_dia(n, n, S["ndiag"], *S["dia"], v, D)}[fmt]
    dev = _Dev(fmt, S[fmt])
    A = mp.Sparse()
    A.format = {"ell": mp.SPARSE_ELL, "jad": mp.SPARSE_JAD, "dia": mp.SPARSE_DIA}[fmt]
    A.maxnzr, A.ndiag = S["maxnzr"], S["ndiag"]
    setattr(A, fmt, dev.s)
    tol = 1e-25 if name == "8_2double" else 1e-90
    for precond in (None, 1.0 / diag):
        note("cg formats", name, fmt, precond is not None)
        x0 = o.set_d(np.zeros(n), D)
        xw, rw = o.cg_csr(n, None, None, None, b, tol, 300, x0, precond=precond, spmv=spmv)
        db, dxx = c.aos(b), c.aos(x0)
        dm = mp.DevBuf(host=precond) if precond is not None else None
        res, it = (C.c_double * 301)(), C.c_int()
        mp.check(c.L.mpres_cg(c.h, n, C.byref(A), db.ptr, tol, 300, dxx.ptr, dm.ptr if dm else None, res,
                              C.byref(it), 32, 256, None, 0, None))
        assert it.value == len(rw) - 1 and list(res[:it.value + 1]) == rw
        same(c.aos_to_host(dxx, n), xw, f"cg {fmt} {name}")


def test_pool_two_streams():
    """The context's workspace pool shared by calls on two streams: a call on another stream
    waits for the pool's last use, and the pool grows only after that use has ended, so
    every result equals its single-stream value"""
    name, n = "8_2double", 1 << 18
    c, o = mp.Context(name), orc(name)  # a fresh context: its pool starts empty
    H = mp.hip()
    sa, sb = C.c_void_p(), C.c_void_p()
    mp.hipchk(H.hipStreamCreate(C.byref(sa)))
    mp.hipchk(H.hipStreamCreate(C.byref(sb)))
    try:
        xd, yd = [gen.uexp(s, n, -20, 20) for s in (41, 42)], [gen.uexp(s, n, -20, 20) for s in (43, 44)]
        ax, ay = [c.array_set_d(v) for v in xd], [c.array_set_d(v) for v in yd]
        r = [c.array(1) for _ in range(3)]
        calls = [(0, 64, 64, sa), (1, 256, 64, sb), (0, 256, 64, sa)]  # grows on the second call
        note("pool two streams", name, n, calls)
        for i, (j, g3, b3, st) in enumerate(calls):
            mp.check(c.L.mpres_dot(c.h, n, C.byref(ax[j]), 1, C.byref(ay[j]), 1, C.byref(r[i]), g3, b3, None, 0, st))
        mp.hipchk(H.hipStreamSynchronize(sa))
        mp.hipchk(H.hipStreamSynchronize(sb))
        for i, (j, g3, b3, _) in enumerate(calls):
            want = o.dot(o.set_d(xd[j], D), o.set_d(yd[j], D), g3, b3, D)
            same(c.array_to_host(r[i], 1), want, f"pool call {i}")
    finally:
        H.hipStreamDestroy(sa)
        H.hipStreamDestroy(sb)
        c.close()


@pytest.mark.parametrize("name", ["8_2double", "32_8double"])
def test_axpy(name):
    c, o = ctx(name), orc(name)
    n = 3000
    X, Y = _operands(o, 31, n // 5)
    alpha = o.set_d([-0.625], D)
    want1 = o.axpy(alpha, X, Y, D)
    want2 = o.axpy_v2(alpha, X, Y, D)
    wants = o.axpy(alpha, X, Y, D, incx=2, incy=-1, n=len(X) // 2)
    for L in lanes_of(c):
        c.set_lanes(mp.K_AXPY, L)
        note("axpy", name, L)
        ax, ay, aa = c.array_from_host(X), c.array_from_host(Y), c.array_from_host(alpha)
        mp.check(c.L.mpres_axpy(c.h, len(X), C.byref(aa), C.byref(ax), 1, C.byref(ay), 1, None))
        same(c.array_to_host(ay, len(X)), want1, f"axpy {name} L={L}")
        ay = c.array_from_host(Y)
        mp.check(c.L.mpres_axpy(c.h, len(X) // 2, C.byref(aa), C.byref(ax), 2, C.byref(ay), -1, None))
        same(c.array_to_host(ay, len(X)), wants, f"axpy strided {name} L={L}")
        dX, dY, dA, dR = c.aos(X), c.aos(Y), c.aos(alpha), c.aos(count=len(X))
        mp.check(c.L.mpres_axpy_v2(c.h, len(X), dA.ptr, dX.ptr, dY.ptr, dR.ptr, None))
        same(c.aos_to_host(dR, len(X)), want2, f"axpy_v2 {name} L={L}")


@pytest.mark.parametrize("name,m,n,b3", [("32_8double", 70, 50, 32), ("32_8double", 33, 64, 5),
                                         ("8_2double", 64, 40, 32), ("16_4double", 20, 90, 16)])
def test_gemv(name, m, n, b3):
    c, o = ctx(name), orc(name)
    A = o.set_d(gen.uexp(41, m * n, -30, 30), D)
    alpha = o.set_d(gen.uexp(42, 1, -30, 30), D)
    beta = o.set_d(gen.uexp(43, 1, -30, 30), D)
    for trans in (0, 1):
        lx, ly = (m, n) if trans else (n, m)
        x = o.set_d(gen.uexp(44, lx, -30, 30), D)
        y = o.set_d(gen.uexp(45, ly, -30, 30), D)
        want = o.gemv(trans, m, n, alpha, A, m, x, beta, y, b3, D)
        for L in lanes_of(c):
            c.set_lanes(mp.K_GEMV, L)
            note("gemv", name, m, n, b3, trans, L)
            aA, ax, ay = c.array_from_host(A), c.array_from_host(x), c.array_from_host(y)
            aa, ab = c.array_from_host(alpha), c.array_from_host(beta)
            mp.check(c.L.mpres_gemv(c.h, mp.TRANS if trans else mp.NO_TRANS, m, n, C.byref(aa), C.byref(aA), m,
                                    C.byref(ax), 1, C.byref(ab), C.byref(ay), 1, b3, None, 0, None))
            same(c.array_to_host(ay, ly), want, f"gemv {name} trans={trans} L={L}")


@pytest.mark.parametrize("name,m,n,k", [("32_8double", 13, 11, 17), ("8_2double", 9, 7, 40), ("16_4double", 5, 6, 7)])
def test_gemm(name, m, n, k):
    c, o = ctx(name), orc(name)
    A = o.set_d(gen.uexp(51, m * k, -30, 30), D)
    B = o.set_d(gen.uexp(52, k * n, -30, 30), D)
    Cm = o.set_d(gen.uexp(53, m * n, -30, 30), D)
    alpha = o.set_d(gen.uexp(54, 1, -30, 30), D)
    beta = o.set_d(gen.uexp(55, 1, -30, 30), D)
    want = o.gemm(m, n, k, alpha, A, m, B, k, beta, Cm, m, D)
    for L in lanes_of(c):
        c.set_lanes(mp.K_GEMM, L)
        note("gemm", name, m, n, k, L)
        aA, aB, aC = c.array_from_host(A), c.array_from_host(B), c.array_from_host(Cm)
        aa, ab = c.array_from_host(alpha), c.array_from_host(beta)
        mp.check(c.L.mpres_gemm(c.h, mp.NO_TRANS, mp.NO_TRANS, m, n, k, C.byref(aa), C.byref(aA), m, C.byref(aB), k,
                                C.byref(ab), C.byref(aC), m, None, 0, None))
        same(c.array_to_host(aC, m * n), want, f"gemm {name} L={L}")
        assert c.L.mpres_gemm(c.h, mp.TRANS, mp.NO_TRANS, m, n, k, C.byref(aa), C.byref(aA), k, C.byref(aB), k,
                              C.byref(ab), C.byref(aC), m, None, 0, None) == -4
    for ta, tb in ((0, 0), (1, 0), (0, 1), (1, 1)):
        lda = k if ta else m
        ldb = n if tb else k
        want2 = o.gemm_v2(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, Cm, m, D)
        for L in lanes_of(c):
            c.set_lanes(mp.K_GEMM, L)
            note("gemm_v2", name, ta, tb, L)
            dA, dB, dC, da, db = c.aos(A), c.aos(B), c.aos(Cm), c.aos(alpha), c.aos(beta)
            mp.check(c.L.mpres_gemm_v2(c.h, mp.TRANS if ta else mp.NO_TRANS, mp.TRANS if tb else mp.NO_TRANS, m, n, k,
                                       da.ptr, dA.ptr, lda, dB.ptr, ldb, db.ptr, dC.ptr, m, None, 0, None))
            same(c.aos_to_host(dC, m * n), want2, f"gemm_v2 {name} {ta}{tb} L={L}")


def _spmv_case(name, m, irp, ja, as_, xs, tprs):
    c, o = ctx(name), orc(name)
    x = o.set_d(xs, D)
    dirp, dja, das, dx = mp.DevBuf(host=irp), mp.DevBuf(host=ja), mp.DevBuf(host=as_), c.aos(x)
    csr = mp.Csr(das.ptr.value, dja.ptr.value, dirp.ptr.value)
    for tpr in tprs:
        want = o.spmv(m, irp, ja, as_, x, tpr, D)
        for L in lanes_of(c):
            c.set_lanes(mp.K_SPMV, L)
            note("spmv", name, m, tpr, L)
            dy = c.aos(count=m)
            mp.check(c.L.mpres_spmv_csrv(c.h, m, C.byref(csr), tpr, dx.ptr, dy.ptr, None))
            same(c.aos_to_host(dy, m), want, f"spmv {name} tpr={tpr} L={L}")


def test_spmv_lf10():
    """tests/sparse/matrices/LF10.mtx, the reference's SpMV fixture."""
    m, n, irp, ja, as_ = gen.read_mtx(os.path.join(ROOT, "tests", "golden", "LF10.mtx"))
    assert (m, n, irp[-1]) == (18, 18, 82)
    _spmv_case("16_4double", m, irp, ja, as_, gen.uexp(61, n, -10, 10), (1, 2, 4, 8, 16, 32))


@pytest.mark.parametrize("name", ["16_4double", "8_2double"])
def test_spmv_powerlaw(name):
    m = 600
    irp, ja, as_ = gen.powerlaw_csr(62, m, kmax=200)
    _spmv_case(name, m, irp, ja, as_, gen.uexp(63, m, -10, 10), (1, 4, 32))


# ---- chains that leave the common case at 16 and 32 moduli ----
# Wide exponent ranges make aligned sums exceed the precision (rounding), and
# exact cancellations make the sum's interval straddle zero (sign estimate by
# mixed-radix comparison). Over 16 and 32 moduli such chains are abandoned by
# the fast kernels and recomputed by the fixup kernels; every lanes variant
# must still agree with the oracle bit for bit.
WIDE = {"16_4double": (-150, 150), "32_8double": (-400, 400)}


def _cancel_pairs(x, y, stride, frac, seed):
    """y[i + stride] = -y[i], x[i + stride] = x[i] for a random subset of i"""
    rng = np.random.default_rng(seed)
    idx = np.nonzero(rng.random(len(x) - stride) < frac)[0]
    idx = idx[((idx // stride) % 2) == 0]
    x[idx + stride] = x[idx]
    y[idx + stride] = -y[idx]
    return x, y


@pytest.mark.parametrize("name", ["16_4double", "32_8double"])
def test_dot_rare(name):
    c, o = ctx(name), orc(name)
    n, g3, b3 = 4000, 16, 32
    lo, hi = WIDE[name]
    xs, ys = _cancel_pairs(gen.uexp(71, n, lo, hi), gen.uexp(72, n, lo, hi), g3 * b3, 0.3, 73)
    # second half: narrow exponents with cancellations only
    xs[n // 2:] = gen.uexp(74, n - n // 2, -10, 10)
    ys[n // 2:] = gen.uexp(75, n - n // 2, -10, 10)
    xs[n // 2:], ys[n // 2:] = _cancel_pairs(xs[n // 2:], ys[n // 2:], g3 * b3, 0.5, 76)
    x, y = o.set_d(xs, D), o.set_d(ys, D)
    ax, ay, r = c.array_from_host(x), c.array_from_host(y), c.array(1)
    dx, dy, dr = c.aos(x), c.aos(y), c.aos(count=1)
    want = o.dot(x, y, g3, b3, D)
    for L in lanes_of(c):
        c.set_lanes(mp.K_DOT, L)
        note("dot_rare", name, L)
        mp.check(c.L.mpres_dot(c.h, n, C.byref(ax), 1, C.byref(ay), 1, C.byref(r), g3, b3, None, 0, None))
        same(c.array_to_host(r, 1), want, f"dot_rare {name} L={L}")
        mp.check(c.L.mpres_dot_v2(c.h, n, dx.ptr, dy.ptr, dr.ptr, g3, b3, None, 0, None))
        same(c.aos_to_host(dr, 1), want, f"dot_v2_rare {name} L={L}")


@pytest.mark.parametrize("name", ["16_4double", "32_8double"])
def test_gemv_rare(name):
    c, o = ctx(name), orc(name)
    m, n, b3 = 24, 96, 8
    lo, hi = WIDE[name]
    As = gen.uexp(81, m * n, lo, hi).reshape(n, m)  # column-major: row j of this array is column j of A
    xs = gen.uexp(82, n, lo, hi)
    # columns j and j + b3 equal with x[j + b3] = -x[j] for some j: exact cancellation in a strided sum
    for j in range(0, n - b3, 2 * b3):
        if j % 3 == 0:
            As[j + b3] = As[j]
            xs[j + b3] = -xs[j]
    A = o.set_d(As.ravel(), D)
    alpha, beta = o.set_d([1.5], D), o.set_d([-0.75], D)
    x = o.set_d(xs, D)
    y = o.set_d(gen.uexp(83, m, lo, hi), D)
    want = o.gemv(0, m, n, alpha, A, m, x, beta, y, b3, D)
    for L in lanes_of(c):
        c.set_lanes(mp.K_GEMV, L)
        note("gemv_rare", name, L)
        aA, ax, ay = c.array_from_host(A), c.array_from_host(x), c.array_from_host(y)
        aa, ab = c.array_from_host(alpha), c.array_from_host(beta)
        mp.check(c.L.mpres_gemv(c.h, mp.NO_TRANS, m, n, C.byref(aa), C.byref(aA), m, C.byref(ax), 1, C.byref(ab),
                                C.byref(ay), 1, b3, None, 0, None))
        same(c.array_to_host(ay, m), want, f"gemv_rare {name} L={L}")


@pytest.mark.parametrize("name", ["16_4double", "32_8double"])
def test_gemm_rare(name):
    c, o = ctx(name), orc(name)
    m, n, k = 9, 8, 24
    lo, hi = WIDE[name]
    As = gen.uexp(91, m * k, lo, hi).reshape(k, m)  # column-major: As[i] is column i of A
    Bs = gen.uexp(92, k * n, lo, hi).reshape(n, k)  # Bs[j] is column j of B
    # A[:, i + 1] = A[:, i] and B[i + 1, j] = -B[i, j]: consecutive terms cancel for even i < 8
    for i in range(0, 8, 2):
        As[i + 1] = As[i]
        Bs[:, i + 1] = -Bs[:, i]
    A, B = o.set_d(As.ravel(), D), o.set_d(Bs.ravel(), D)
    Cm = o.set_d(gen.uexp(93, m * n, lo, hi), D)
    alpha, beta = o.set_d([0.5], D), o.set_d([3.0], D)
    want = o.gemm(m, n, k, alpha, A, m, B, k, beta, Cm, m, D)
    want2 = o.gemm_v2(0, 0, m, n, k, alpha, A, m, B, k, beta, Cm, m, D)
    for L in lanes_of(c):
        c.set_lanes(mp.K_GEMM, L)
        note("gemm_rare", name, L)
        aA, aB, aC = c.array_from_host(A), c.array_from_host(B), c.array_from_host(Cm)
        aa, ab = c.array_from_host(alpha), c.array_from_host(beta)
        mp.check(c.L.mpres_gemm(c.h, mp.NO_TRANS, mp.NO_TRANS, m, n, k, C.byref(aa), C.byref(aA), m, C.byref(aB), k,
                                C.byref(ab), C.byref(aC), m, None, 0, None))
        same(c.array_to_host(aC, m * n), want, f"gemm_rare {name} L={L}")
        dA, dB, dC, da, db = c.aos(A), c.aos(B), c.aos(Cm), c.aos(alpha), c.aos(beta)
        mp.check(c.L.mpres_gemm_v2(c.h, mp.NO_TRANS, mp.NO_TRANS, m, n, k, da.ptr, dA.ptr, m, dB.ptr, k, db.ptr,
                                   dC.ptr, m, None, 0, None))
        same(c.aos_to_host(dC, m * n), want2, f"gemm_v2_rare {name} L={L}")


@pytest.mark.parametrize("name", ["16_4double", "32_8double"])
def test_spmv_rare(name):
    m = 300
    irp, ja, as_ = gen.powerlaw_csr(101, m, kmax=100)
    lo, hi = WIDE[name]
    xs = gen.uexp(102, m, lo, hi)
    # rows whose first two entries hit one column with opposite values cancel exactly
    for r in range(0, m, 3):
        s, e = irp[r], irp[r + 1]
        if e - s >= 2:
            ja[s + 1] = ja[s]
            as_[s + 1] = -as_[s]
    _spmv_case(name, m, irp, ja, as_, xs, (1, 2, 8))


# ---- precision form on INT8 MFMA (mpres_gemm_mfma) ----
def _log2m(c):
    v = [C.c_int() for _ in range(5)]
    mp.check(c.L.mpres_ctx_info(c.h, *[C.byref(x) for x in v]))
    return v[4].value


def mfma_reference(o, log2m, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc, stats=None):
    """The precision form restated with exact integers and the oracle's primitives.
    Per row of op(A) / column of op(B), over its nonzero elements: emin, the highest bit
    top = max(eval_up.exp + 2 + log2m + exp); the high band's exponent eh = max(emin,
    top - half) holds the elements with exponent >= eh exactly; the others form the low
    band, exponent emin raised by its excess over the budget (truncating). Per entry and
    band pair (low x low, low x high, high x low, high x high; pairs with an empty band
    skipped) the exact signed sum of the aligned products, evaluated (oracle, with the
    2^K0 scaling), rounded, added in that order (mp_add), then the epilogue (mp_mul,
    mp_add). Returns the result and the high x high sums (rounded)."""
    mm = [int(v) for v in o.moduli()]
    lk = (max(k, 1) - 1).bit_length()
    half = (log2m - 2 - lk) // 2
    NONE = -(1 << 30)

    def opa(row, i):
        return A[i + lda * row] if ta else A[row + lda * i]

    def opb(i, col):
        return B[col + ldb * i] if tb else B[i + ldb * col]

    def top_of(r):
        return int(r["eval"][1]["exp"]) + 2 + log2m + int(r["exp"])

    def bands(elems):
        """[(low exponent, low bits), (high exponent, high bits)] of a row or column, and the
        band (0 low, 1 high) of each element"""
        nz = [r for r in elems if float(r["eval"][1]["frac"]) != 0]
        if not nz:
            return [(0, NONE), (0, NONE)], [1] * len(elems)

        def band(emin, t):
            span = t - emin
            return emin + max(0, span - half), min(span, half)

        emin = min(int(r["exp"]) for r in nz)
        eh, bh = band(emin, max(top_of(r) for r in nz))
        low = [r for r in nz if int(r["exp"]) < eh]
        el, bl = band(emin, max(top_of(r) for r in low)) if low else (0, NONE)
        return [(el, bl), (eh, bh)], [1 if int(r["exp"]) >= eh else 0 for r in elems]

    def aligned(r, eff):
        X = exact.crt(r["digits"], mm)
        d = int(r["exp"]) - eff
        v = X << d if d >= 0 else X >> (-d)
        return -v if int(r["sign"]) else v

    def side(elems_of, cnt):
        info = [bands(elems_of(j)) for j in range(cnt)]
        vals = []
        for j in range(cnt):
            (bd, which), elems = info[j], elems_of(j)
            v = [[0] * k, [0] * k]
            for i, r in enumerate(elems):
                v[which[i]][i] = aligned(r, bd[which[i]][0])
            vals.append(v)
        return [x[0] for x in info], vals

    ba_, av = side(lambda row: [opa(row, i) for i in range(k)], m)
    bb_, bv = side(lambda col: [opb(i, col) for i in range(k)], n)
    pairs = []
    for p in range(4):
        pa, pb = p >> 1, p & 1
        sums = o.empty(m * n)
        scaled = np.zeros((m * n, len(mm)), np.int32)
        k0s = np.zeros(m * n, np.int64)
        live = np.zeros(m * n, bool)
        for col in range(n):
            for row in range(m):
                e = row + m * col
                (ea, ba), (eb, bb) = ba_[row][pa], bb_[col][pb]
                live[e] = p == 3 or ((pa == 1 or ba != NONE) and (pb == 1 or bb != NONE))
                S = sum(av[row][pa][i] * bv[col][pb][i] for i in range(k))
                rec = sums[e]
                rec["digits"] = [abs(S) % q for q in mm]
                rec["sign"] = int(S < 0)
                rec["exp"] = ea + eb
                # evaluation of |S| 2^K0 (|S| 2^K0 < M/2), exponents lowered by K0
                k0 = min(max(log2m - 1 - (max(ba, 0) + max(bb, 0) + lk), 0), log2m - 1)
                scaled[e] = [(abs(S) << k0) % q for q in mm]
                k0s[e] = k0
        lo, up = o.eval(scaled, fast=False)
        lo["exp"] = np.where(lo["frac"] != 0, lo["exp"] - k0s, lo["exp"])
        up["exp"] = np.where(up["frac"] != 0, up["exp"] - k0s, up["exp"])
        sums["eval"][:, 0] = lo
        sums["eval"][:, 1] = up
        pairs.append((o.round(sums), live))
    total = o.empty(m * n)
    for e in range(m * n):
        acc = None
        for sums, live in pairs:
            if live[e]:
                acc = sums[e:e + 1] if acc is None else o.add(acc, sums[e:e + 1], D)
        total[e] = acc[0]
    def rounded(e, row, col):
        s2 = o.mul(total[e:e + 1], alpha, D)
        c2 = o.mul(Cm[col * ldc + row:col * ldc + row + 1], beta, D)
        return o.add(s2, c2, D)[0]

    if len(mm) == 8:
        return mfma_exact_epilogue(o, mm, log2m, m, n, k, ba_, bb_, av, bv, pairs, alpha, beta, Cm, ldc,
                                   rounded, stats), total
    want = np.array(Cm, copy=True)
    for col in range(n):
        for row in range(m):
            want[col * ldc + row] = rounded(row + m * col, row, col)
    return want, total


MF_XW_BITS = 32 * 48  # the widest exact accumulation of the 8-moduli finish


def mfma_exact_epilogue(o, mm, log2m, m, n, k, ba_, bb_, av, bv, pairs, alpha, beta, Cm, ldc, rounded, stats=None):
    """The 8-moduli finish: alpha times the exact band-pair sums plus beta times C, formed
    exactly and rounded once to MP_PRECISION bits to nearest (ties to even); the significand
    without trailing zeros, its residues and the oracle's full evaluation. An entry whose
    terms span more than MF_XW_BITS (exponent span + 2 (log2 M + 1) + 2 bits) takes the
    rounded finish (rounded(e, row, col)). stats, when given, counts the entries by the
    accumulator they need: [common (768 bits), wide (1536 bits), rounded]."""
    prec = o.mp_consts()[0]
    M = exact.moduli_product(mm)

    def val(r):
        return (-1) ** int(r["sign"]) * exact.crt(r["digits"], mm), int(r["exp"])

    xa, ea = val(alpha[0])
    xb, eb = val(beta[0])
    want = np.array(Cm, copy=True)
    recs, where = [], []
    for col in range(n):
        for row in range(m):
            e = row + m * col
            terms = []
            for p, (_, live) in enumerate(pairs):
                pa, pb = p >> 1, p & 1
                if not live[e] or xa == 0:
                    continue
                S = sum(av[row][pa][i] * bv[col][pb][i] for i in range(k))
                if S:
                    terms.append((S * xa, ba_[row][pa][0] + bb_[col][pb][0] + ea))
            xc, ec = val(Cm[col * ldc + row])
            if xb and xc:
                terms.append((xc * xb, ec + eb))
            r = o.empty(1)[0]
            if terms:
                emin = min(t[1] for t in terms)
                need = max(t[1] for t in terms) - emin + 2 * (log2m + 1) + 2
                if stats is not None:
                    stats[0 if need <= 32 * 24 else 1 if need <= MF_XW_BITS else 2] += 1
                if need > MF_XW_BITS:
                    want[col * ldc + row] = rounded(e, row, col)
                    continue
                T = sum(v << (x - emin) for v, x in terms)
                if T:
                    q, sh = abs(T), 0
                    if q.bit_length() > prec:
                        sh = q.bit_length() - prec
                        low, q = q & ((1 << sh) - 1), q >> sh
                        half = 1 << (sh - 1)
                        if low > half or (low == half and q & 1):
                            q += 1
                    tz = (q & -q).bit_length() - 1
                    q >>= tz
                    assert q < M
                    r["digits"] = [q % v for v in mm]
                    r["sign"] = int(T < 0)
                    r["exp"] = emin + sh + tz
            recs.append(r)
            where.append(col * ldc + row)
    recs = np.array(recs)
    lo, up = o.eval(np.ascontiguousarray(recs["digits"]), fast=False, policy=D)
    recs["eval"][:, 0] = lo
    recs["eval"][:, 1] = up
    for r, w in zip(recs, where):
        want[w] = r
    return want


@pytest.mark.parametrize("name,lo,hi", [("8_2double", -8, 8), ("8_2double", -60, 60), ("8_2double", -150, 150),
                                        ("16_4double", -20, 20), ("32_8double", -60, 60), ("16_32bit", -30, 30)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 1)])
def test_gemm_mfma(name, lo, hi, ta, tb):
    c, o = ctx(name), orc(name)
    m, n, k = 37, 29, 45
    lda = k if ta else m
    ldb = n if tb else k
    As = gen.uexp(121, m * k, lo, hi)
    Bs = gen.uexp(122, k * n, lo, hi)
    As[5] = 0.0
    Bs[7] = -0.0
    A, B = o.set_d(As, D), o.set_d(Bs, D)
    Cm = o.set_d(gen.uexp(123, m * n, lo, hi), D)
    alpha, beta = o.set_d([1.25], D), o.set_d([-0.5], D)
    want, sums = mfma_reference(o, _log2m(c), ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, Cm, m)
    for L in lanes_of(c):
        c.set_lanes(mp.K_MFMA, L)
        note("gemm_mfma", name, lo, hi, ta, tb, L)
        aA, aB, aC = c.array_from_host(A), c.array_from_host(B), c.array_from_host(Cm)
        aa, ab = c.array_from_host(alpha), c.array_from_host(beta)
        mp.check(c.L.mpres_gemm_mfma(c.h, mp.TRANS if ta else mp.NO_TRANS, mp.TRANS if tb else mp.NO_TRANS, m, n, k,
                                     C.byref(aa), C.byref(aA), lda, C.byref(aB), ldb, C.byref(ab), C.byref(aC), m,
                                     None, 0, None))
        same(c.array_to_host(aC, m * n), want, f"gemm_mfma {name} [{lo},{hi}] {ta}{tb} L={L}")
    # accuracy of the sums against the exact product: exact band sums, rounded and added (a few
    # roundings); e in [-150, 150] at 8 moduli exceeds two bands, so the low band truncates there
    mm = [int(v) for v in o.moduli()]
    prec = o.mp_consts()[0]
    worst = 0.0
    for col in range(0, n, 7):
        for row in range(0, m, 5):
            ex = sum(exact.mp_value(A[(i + lda * row) if ta else (row + lda * i)], mm) *
                     exact.mp_value(B[(col + ldb * i) if tb else (i + ldb * col)], mm) for i in range(k))
            got = exact.mp_value(sums[row + m * col], mm)
            if ex != 0:
                worst = max(worst, float(abs(got - ex) / abs(ex)))
    if hi <= 60:
        assert worst <= 2.0 ** (-prec + 3), worst


def test_gemm_mfma_wide_finish():
    """8 moduli: C entries far above alpha op(A) op(B) make terms that span more than the
    common accumulator (768 bits: the wide exact finish over the list of such entries) or
    more than 1536 bits (the rounded finish); each kind is present"""
    name = "8_2double"
    c, o = ctx(name), orc(name)
    m, n, k = 23, 19, 17
    As, Bs = gen.uexp(131, m * k, -8, 8), gen.uexp(132, k * n, -8, 8)
    Cs = gen.uexp(133, m * n, -8, 8)
    Cs[1::3] *= 2.0 ** 400   # about 600 bits above the products: wide
    Cs[2::3] *= 2.0 ** 900   # about 1100 bits above: rounded
    Cs[7] = 0.0
    A, B, Cm = o.set_d(As, D), o.set_d(Bs, D), o.set_d(Cs, D)
    alpha, beta = o.set_d([2.0 ** -200 * 1.375], D), o.set_d([0.75], D)
    stats = [0, 0, 0]
    want, _ = mfma_reference(o, _log2m(c), 0, 0, m, n, k, alpha, A, m, B, k, beta, Cm, m, stats)
    assert all(v > 0 for v in stats), stats
    note("gemm_mfma wide finish", stats)
    aA, aB, aC = c.array_from_host(A), c.array_from_host(B), c.array_from_host(Cm)
    aa, ab = c.array_from_host(alpha), c.array_from_host(beta)
    for rep in range(2):  # the second call: the list count starts from zero again
        aC = c.array_from_host(Cm)
        mp.check(c.L.mpres_gemm_mfma(c.h, mp.NO_TRANS, mp.NO_TRANS, m, n, k, C.byref(aa), C.byref(aA), m,
                                     C.byref(aB), k, C.byref(ab), C.byref(aC), m, None, 0, None))
        same(c.array_to_host(aC, m * n), want, f"gemm_mfma wide finish call {rep}")


def _batch_case(c, o, name, seed0):
    """a mixed batch on the device: dots and gemvs of several shapes, strides and orders,
    plus a skipped dot; returns (problems, device arrays, [(kind, array, count, expected, reset)])"""
    lo, hi = (-20, 20) if name.startswith("8_") else (-60, 60)
    seed = [seed0]
    keep = []

    def rec(cnt):
        seed[0] += 1
        return o.set_d(gen.uexp(seed[0], cnt, lo, hi), D)

    def dev(recs):
        a = c.array_from_host(recs)
        keep.append(a)
        return a

    probs, checks = [], []
    for n, g3, b3, incx, incy in [(1, 1, 1, 1, 1), (777, 2, 3, 1, 1), (3000, 4, 64, 1, 1), (257, 3, 5, -1, 2),
                                  (4100, 1, 32, 2, -1), (64, 8, 8, 1, 1), (8192, 256, 64, 1, 1)]:
        xr, yr = rec(1 + (n - 1) * abs(incx)), rec(1 + (n - 1) * abs(incy))
        x, y, r = dev(xr), dev(yr), c.array(1)
        probs.append(mp.dot_problem(n, x, y, r, g3, b3, incx, incy))
        checks.append(("dot", r, 1, o.dot(xr, yr, g3, b3, D, incx=incx, incy=incy, n=n), None))
    for m, n, trans, b3, incx, incy in [(5, 7, 0, 1, 1, 1), (33, 20, 0, 3, 1, 1), (33, 20, 1, 8, 1, 1),
                                        (130, 90, 0, 2, -1, 2), (40, 100, 1, 1, 2, -1), (1, 1, 0, 4, 1, 1),
                                        (300, 257, 0, 32, 1, 1)]:
        lda = m + 3
        lx, ly = (m, n) if trans else (n, m)
        Ar, xr, yr, ar, br = rec(lda * n), rec(1 + (lx - 1) * abs(incx)), rec(1 + (ly - 1) * abs(incy)), rec(1), rec(1)
        A, x, y, al, be = dev(Ar), dev(xr), dev(yr), dev(ar), dev(br)
        probs.append(mp.gemv_problem(mp.TRANS if trans else mp.NO_TRANS, m, n, al, A, lda, x, be, y, b3, incx, incy))
        checks.append(("gemv", y, len(yr), o.gemv(trans, m, n, ar, Ar, lda, xr, br, yr, b3, D, incx=incx, incy=incy),
                       yr))
    zr = rec(1)
    z, zout = dev(rec(3)), dev(zr)
    probs.append(mp.dot_problem(0, z, z, zout, 2, 2))
    checks.append(("dot", zout, 1, zr, None))
    return probs, keep, checks


@pytest.mark.parametrize("name", ["16_4double", "8_2double", "32_8double"])
def test_batch(name):
    """mpres_batch_run: every problem bit-identical to its single call (oracle, policy D), for
    every lanes variant and several geometries; the graph replayed three times, the updated
    gemv outputs restored between replays, every replay checked"""
    c, o = ctx(name), orc(name)
    probs, keep, checks = _batch_case(c, o, name, 600)
    geoms = [dict(lanes=L, wg_per_cu=2, gemv_rows=64) for L in lanes_of(c)]
    geoms += [dict(lanes=lanes_of(c)[-1], wg_per_cu=1, gemv_rows=16), dict(lanes=0, wg_per_cu=4, gemv_rows=1000)]
    for g in geoms:
        note("batch", name, g)
        b = mp.Batch(c, probs, g)
        info = b.info()
        assert info["mp_fmas"] == sum(p.n if p.op == mp.BATCH_DOT else p.m * p.n for p in probs)
        for rep in range(3):
            for kind, arr, cnt, want, reset in checks:
                if reset is not None:
                    mp.check(c.L.mpres_array_host2device(c.h, C.byref(arr), reset.ctypes.data_as(C.c_void_p), cnt,
                                                         None))
            mp.synchronize()
            b.run()
            mp.synchronize()
            for i, (kind, arr, cnt, want, reset) in enumerate(checks):
                same(c.array_to_host(arr, cnt), want, f"batch {name} {g} replay {rep} problem {i} ({kind})")
        b.close()


def test_batch_errors():
    c = ctx("16_4double")
    x = c.array(4)
    bad = mp.dot_problem(4, x, x, x, 0, 4)
    with pytest.raises(mp.MpresError, match="problem 0"):
        mp.Batch(c, [bad])
    with pytest.raises(mp.MpresError):
        mp.Batch(c, [mp.dot_problem(4, x, x, x, 1, 4)], dict(lanes=3, wg_per_cu=1, gemv_rows=8))
    empty = mp.Batch(c, [])
    assert empty.info()["items"] == 0
    empty.run()
    empty.close()


def test_dot_headline_config():
    """The CPU-runnable headline case: mp_dot of two vectors of n = 65536, u*2^e with e in
    [-20, 20], 8 moduli (106 bits), order 256 x 64, inputs converted on the GPU: bit-exact
    against the oracle, and its error against the exact sum within what the reference's
    rounded chain allows (measured 1.08e-31, about 9 units of 2^-106)"""
    c, o = ctx("8_2double"), orc("8_2double")
    n = 65536
    hx, hy = gen.uexp(1, n, -20, 20), gen.uexp(2, n, -20, 20)
    x, y = o.set_d(hx, D), o.set_d(hy, D)
    want = o.dot(x, y, 256, 64, D)
    ax, ay, r = c.array_set_d(hx), c.array_set_d(hy), c.array(1)
    c.set_lanes(mp.K_DOT, 0)
    note("dot headline", n)
    same(c.array_to_host(ax, n), x, "set_d x")
    mp.check(c.L.mpres_dot(c.h, n, C.byref(ax), 1, C.byref(ay), 1, C.byref(r), 256, 64, None, 0, None))
    got = c.array_to_host(r, 1)
    same(got, want, "dot 65536")
    m = [int(v) for v in o.moduli()]
    ex = sum(Fraction(float(a)) * Fraction(float(b)) for a, b in zip(hx, hy))
    err = abs(exact.mp_value(got[0], m) - ex) / abs(ex)
    assert err < 2.0 ** -95, float(err)


def test_gemm_rejects_misaligned_digits():
    """both gemm entry points read residues as 16-byte vectors: digits of A or B off a
    16-byte boundary return MPRES_ERR_ARG before anything is launched"""
    c = ctx("16_4double")
    A, B, Cm = c.array_set_d(gen.uexp(141, 4, -3, 3)), c.array_set_d(gen.uexp(142, 4, -3, 3)), \
        c.array_set_d(gen.uexp(143, 4, -3, 3))
    al, be = c.array_set_d([1.5]), c.array_set_d([0.5])
    off = mp.MpArray.from_buffer_copy(A)
    off.digits = A.digits + 4  # still inside A's digits; never read
    for fn in (c.L.mpres_gemm, c.L.mpres_gemm_mfma):
        for a, b in ((off, B), (B, off)):
            rc = fn(c.h, mp.NO_TRANS, mp.NO_TRANS, 2, 2, 2, C.byref(al), C.byref(a), 2, C.byref(b), 2,
                    C.byref(be), C.byref(Cm), 2, None, 0, None)
            assert rc == -1, rc  # MPRES_ERR_ARG
        rc = fn(c.h, mp.NO_TRANS, mp.NO_TRANS, 2, 2, 2, C.byref(al), C.byref(A), 2, C.byref(B), 2,
                C.byref(be), C.byref(Cm), 2, None, 0, None)
        assert rc == 0, rc
    mp.synchronize()


@pytest.mark.parametrize("name", ["16_4double", "32_8double"])
def test_gemm_long_k(name):
    """k beyond one INT32 accumulation of the MFMA residue products (k > 16384): the
    chunked kernel folds its accumulators every 8192 k; bit-exact against the oracle"""
    c, o = ctx(name), orc(name)
    m, n, k = 5, 3, 16500
    A = o.set_d(gen.uexp(131, m * k, -30, 30), D)
    B = o.set_d(gen.uexp(132, k * n, -30, 30), D)
    Cm = o.set_d(gen.uexp(133, m * n, -30, 30), D)
    alpha, beta = o.set_d([0.75], D), o.set_d([-2.0], D)
    want = o.gemm(m, n, k, alpha, A, m, B, k, beta, Cm, m, D)
    c.set_lanes(mp.K_GEMM, 0)
    note("gemm long k", name, m, n, k)
    aA, aB, aC = c.array_from_host(A), c.array_from_host(B), c.array_from_host(Cm)
    aa, ab = c.array_from_host(alpha), c.array_from_host(beta)
    mp.check(c.L.mpres_gemm(c.h, mp.NO_TRANS, mp.NO_TRANS, m, n, k, C.byref(aa), C.byref(aA), m, C.byref(aB), k,
                            C.byref(ab), C.byref(aC), m, None, 0, None))
    same(c.array_to_host(aC, m * n), want, f"gemm {name} k={k}")
