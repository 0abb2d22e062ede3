"""The oracle's restatement of the rest of the BLAS (src/blas/*.cuh, src/blas/v2/*.cuh) pinned
without the restatement itself: on small-integer operands every product and sum is exact, so
each routine must return the exact closed form of its definition (which catches a dropped or
doubled term, a wrong index, increment, triangle or transposition); the reductions' fold and
tree shape is pinned to the oracle's C dot, whose order is the reference's, and the maximum
to the exact largest magnitude."""
from fractions import Fraction

import numpy as np
import pytest

import exact
import gen
import pyoracle as po

D = po.POLICY_D
SETS = ["8_2double", "32_8double"]


def ints(seed, count, lo=-40, hi=40):
    return np.random.default_rng(seed).integers(lo, hi + 1, count).astype(np.float64)


def val(o, recs):
    m = [int(v) for v in o.moduli()]
    return [exact.mp_value(r, m) for r in recs]


def bidx(i, n, inc):
    return i * inc if inc > 0 else (-n + i + 1) * inc


@pytest.fixture(scope="module", params=SETS)
def o(request):
    return po.Oracle(request.param)


def test_scal(o):
    n, incx = 30, 3
    xv = ints(1, n * incx)
    x = o.set_d(xv, D)
    a = o.set_d([-3.0], D)
    got = val(o, o.scal(n, a, x, incx))
    want = [Fraction(v) for v in xv]
    for i in range(n):
        want[i * incx] *= -3
    assert got == want
    assert o.scal(n, a, x, 0).tobytes() == x.tobytes()  # incx <= 0: untouched


@pytest.mark.parametrize("incx,incy", [(1, 1), (2, -1), (-3, 2)])
def test_rot(o, incx, incy):
    n = 25
    xv, yv = ints(2, n * abs(incx)), ints(3, n * abs(incy))
    x, y = o.set_d(xv, D), o.set_d(yv, D)
    c, s = o.set_d([3.0], D), o.set_d([-2.0], D)
    gx, gy, b1, b2 = o.rot(n, x, incx, y, incy, c, s)
    wx, wy = [Fraction(v) for v in xv], [Fraction(v) for v in yv]
    # the c scaling covers elements 0 .. n-1 whatever the increments are
    for j in range(n):
        wx[j] *= 3
        wy[j] *= 3
    for i in range(n):
        ix, iy = bidx(i, n, incx), bidx(i, n, incy)
        wx[ix] += -2 * Fraction(yv[iy])
        wy[iy] -= -2 * Fraction(xv[ix])
    assert val(o, gx) == wx and val(o, gy) == wy
    assert val(o, b1) == [-2 * Fraction(xv[bidx(i, n, incx)]) for i in range(n)]
    assert val(o, b2) == [-2 * Fraction(yv[bidx(i, n, incy)]) for i in range(n)]


def test_waxpby_and_axpy_dot(o):
    n = 40
    xv, yv, wv = ints(4, 2 * n), ints(5, n), ints(6, 3 * n)
    x, y, w = o.set_d(xv, D), o.set_d(yv, D), o.set_d(wv, D)
    al, be = o.set_d([5.0], D), o.set_d([-7.0], D)
    gw, buf = o.waxpby(n, al, x, 2, be, y, -1, w, 3)
    want = [Fraction(v) for v in wv]
    for i in range(n):
        want[3 * i] = 5 * Fraction(xv[2 * i]) - 7 * Fraction(yv[n - 1 - i])
    assert val(o, gw) == want
    # axpy_dot: w = w - alpha v, r = u . w
    u, v = o.set_d(ints(7, n), D), o.set_d(ints(8, n), D)
    gw, r = o.axpy_dot(n, al, w, 3, v, 1, u, -1, 16, 8)
    uv, vv = val(o, u), val(o, v)
    want = [Fraction(q) for q in wv]
    for i in range(n):
        want[3 * i] -= 5 * vv[i]
    assert val(o, gw) == want
    assert val(o, r)[0] == sum(uv[n - 1 - i] * want[3 * i] for i in range(n))


@pytest.mark.parametrize("g,b", [(4, 32), (3, 5), (64, 64), (1, 1)])
def test_fold_tree_is_the_dot_order(o, g, b):
    """the oracle's reduction shape with mp_add over products equals its C dot bit for bit"""
    n = 1500
    x = o.set_d(gen.uexp(9, n, -30, 30), D)
    y = o.set_d(gen.uexp(10, n, -30, 30), D)
    prods = o.mul(x, y)
    p = o._fold_tree(lambda i: prods[i], n, g, b, o.add)
    r = o._fold_tree(lambda i: p[i], g, 1, b, o.add)
    assert r.tobytes() == o.dot(x, y, g, b, D).tobytes()


@pytest.mark.parametrize("g,b", [(4, 32), (3, 5), (300, 64)])
def test_asum_and_norm_inf(o, g, b):
    n, incx = 700, 2
    xv = ints(11, n * incx, -1000, 1000)
    xv[::7] = 0.0
    x = o.set_d(xv, D)
    xs = [abs(Fraction(v)) for v in xv[::incx][:n]]
    r = o.asum(n, x, incx, g, b)
    assert r["sign"][0] == 0 and val(o, r)[0] == sum(xs)
    m = o.norm_inf(n, x, incx, g, b)
    assert m["sign"][0] == 0 and val(o, m)[0] == max(xs)
    # the record is an input's with sign 0
    cand = x[::incx][:n].copy()
    cand["sign"] = 0
    assert any(m[0].tobytes() == c.tobytes() for c in cand)
    # general data: the sum within the precision of the exact one
    x = o.set_d(gen.uexp(12, n, -20, 20), D)
    got = val(o, o.asum(n, x, 1, g, b))[0]
    ex = sum(abs(v) for v in val(o, x))
    assert abs(got - ex) <= ex * Fraction(2) ** -90


def test_ge_add_acc_scale(o):
    m, n, lda, ldb, ldc = 7, 6, 9, 8, 10
    Av, Bv, Cv = ints(13, lda * n), ints(14, ldb * n), ints(15, ldc * n)
    A, B, Cm = o.set_d(Av, D), o.set_d(Bv, D), o.set_d(Cv, D)
    al, be = o.set_d([2.0], D), o.set_d([-3.0], D)
    gc, buf = o.ge_add(m, n, al, A, lda, be, B, ldb, Cm, ldc)
    want = [Fraction(v) for v in Cv]
    for j in range(n):
        for i in range(m):
            want[i + j * ldc] = 2 * Fraction(Av[i + j * lda]) - 3 * Fraction(Bv[i + j * ldb])
    assert val(o, gc) == want
    assert val(o, buf) == [2 * Fraction(Av[i + j * lda]) for j in range(n) for i in range(m)]
    gb, _ = o.ge_acc(m, n, al, A, lda, be, B, ldb)
    want = [Fraction(v) for v in Bv]
    for j in range(n):
        for i in range(m):
            want[i + j * ldb] = 2 * Fraction(Av[i + j * lda]) - 3 * Fraction(Bv[i + j * ldb])
    assert val(o, gb) == want
    # diagonal scaling: right A D (column j times D_j), left D A (row i times D_i)
    Dv = ints(16, 2 * max(m, n), 1, 9)
    Dm = o.set_d(Dv, D)
    for right, incd in ((True, 2), (False, -1), (True, -2)):
        cnt = n if right else m
        got = val(o, o.ge_diag_scale(right, m, n, Dm, incd, A, lda))
        want = [Fraction(v) for v in Av]
        for j in range(n):
            for i in range(m):
                d = Dv[bidx(j if right else i, cnt, incd)]
                want[i + j * lda] *= int(d)
        assert got == want
    got = val(o, o.ge_lr_scale(m, n, Dm, 1, Dm, 2, A, lda))
    want = [Fraction(v) for v in Av]
    for j in range(n):
        for i in range(m):
            want[i + j * lda] *= int(Dv[i]) * int(Dv[2 * j])
    assert got == want


@pytest.mark.parametrize("b", [4, 7, 64])
def test_ge_norm(o, b):
    m, n, lda = 13, 9, 15
    Av = ints(17, lda * n)
    A = o.set_d(Av, D)
    col = [sum(abs(Fraction(Av[i + j * lda])) for i in range(m)) for j in range(n)]
    row = [sum(abs(Fraction(Av[i + j * lda])) for j in range(n)) for i in range(m)]
    r, sums = o.ge_norm(True, m, n, A, lda, 3, b)
    assert val(o, sums) == col and val(o, r)[0] == max(col) and all(sums["sign"] == 0)
    r, sums = o.ge_norm(False, m, n, A, lda, 3, b)
    assert val(o, sums) == row and val(o, r)[0] == max(row)


@pytest.mark.parametrize("incx,incy", [(1, 1), (-2, 3)])
def test_ger(o, incx, incy):
    m, n, lda = 8, 5, 11
    xv, yv, Av = ints(18, m * abs(incx)), ints(19, n * abs(incy)), ints(20, lda * n)
    x, y, A = o.set_d(xv, D), o.set_d(yv, D), o.set_d(Av, D)
    al = o.set_d([-4.0], D)
    ga, b1, b2 = o.ger(m, n, al, x, incx, y, incy, A, lda)
    want = [Fraction(v) for v in Av]
    for j in range(n):
        for i in range(m):
            want[i + j * lda] += Fraction(xv[bidx(i, m, incx)]) * -4 * Fraction(yv[bidx(j, n, incy)])
    assert val(o, ga) == want
    # the v2 form: same values
    ga2 = o.ger_v2(m, n, al, x, incx, y, incy, A, lda)
    assert val(o, ga2) == want


def test_v2_scal_syr_syr2(o):
    n, lda = 9, 10
    xv, yv, Av = ints(21, 2 * n), ints(22, n), ints(23, lda * n)
    x, y, A = o.set_d(xv, D), o.set_d(yv, D), o.set_d(Av, D)
    al = o.set_d([3.0], D)
    assert val(o, o.scal_v2(al, x)) == [3 * Fraction(v) for v in xv]
    for upper in (True, False):
        got = val(o, o.syr_v2(upper, n, al, x, -2, A, lda))
        got2 = val(o, o.syr2_v2(upper, n, al, x, 2, y, 1, A, lda))
        want, want2 = [Fraction(v) for v in Av], [Fraction(v) for v in Av]
        for j in range(n):
            for i in range(n):
                if (i <= j) if upper else (j <= i):
                    xi, xj = Fraction(xv[bidx(i, n, -2)]), Fraction(xv[bidx(j, n, -2)])
                    want[i + j * lda] += 3 * xi * xj
                    want2[i + j * lda] += 3 * (Fraction(xv[2 * i]) * Fraction(yv[j]) +
                                               Fraction(yv[i]) * Fraction(xv[2 * j]))
        assert got == want and got2 == want2


def test_v2_symv_syrk(o):
    n, lda, k = 8, 9, 5
    Sv = ints(24, n * n)
    S = np.array(Sv).reshape(n, n)
    S = S + S.T  # symmetric
    for upper in (True, False):
        # A holds the symmetric matrix on the uplo triangle and garbage on the other
        Av = ints(25, lda * n, 500, 900)
        for j in range(n):
            for i in range(n):
                if (i <= j) if upper else (j <= i):
                    Av[i + j * lda] = S[i, j]
        A = o.set_d(Av, D)
        xv, yv = ints(26, 3 * n), ints(27, n)
        x, y = o.set_d(xv, D), o.set_d(yv, D)
        al, be = o.set_d([2.0], D), o.set_d([-1.0], D)
        got = val(o, o.symv_v2(upper, n, al, A, lda, x, 3, be, y, -1))
        want = [Fraction(v) for v in yv]
        for r in range(n):
            want[n - 1 - r] = -Fraction(yv[n - 1 - r]) + 2 * sum(int(S[r, c]) * Fraction(xv[3 * c]) for c in range(n))
        assert got == want
    # syrk: C = (op(A) op(A)^T) alpha + C beta on the triangle; only TRANS reads A^T
    Bv = ints(28, 12 * 12)
    B = o.set_d(Bv, D)
    Cv = ints(29, 10 * n)
    Cm = o.set_d(Cv, D)
    al, be = o.set_d([3.0], D), o.set_d([2.0], D)
    for upper in (True, False):
        for trans in (False, True):
            ldb = 12
            got = val(o, o.syrk_v2(upper, trans, n, k, al, B, ldb, be, Cm, 10))
            want = [Fraction(v) for v in Cv]
            opA = (lambda r, i: Bv[i + ldb * r]) if trans else (lambda r, i: Bv[r + ldb * i])
            for c in range(n):
                for r in range(n):
                    if (r <= c) if upper else (c <= r):
                        s = sum(Fraction(opA(r, i)) * Fraction(opA(c, i)) for i in range(k))
                        want[r + c * 10] = 3 * s + 2 * Fraction(Cv[r + c * 10])
            assert got == want
