"""Pins the oracle's BLAS compositions (dot, axpy, gemv, gemm, SpMV).

Two independent checks per operation:
* value: at 32 moduli (424 bits) and 16 moduli (212 bits) with narrow
  exponents nothing rounds, so every result must equal the exact rational
  value of the BLAS formula (computed here from big-integer CRT values). This
  catches a dropped term, a wrong sign, index, stride, leading dimension or a
  transposed operand.
* order: at 8 moduli (106 bits) almost every accumulation rounds, so the bits
  depend on the summation order. The order is recomposed here from the
  oracle's element-wise primitives (mp_mul, mp_mul_d, mp_add, v1 vector add)
  following the reference kernels cited at each helper, and must match the
  oracle's composed routine bit for bit.
"""
from fractions import Fraction

import numpy as np
import pytest

import exact
import gen
import pyoracle as po

D = po.POLICY_D


def bidx(k, n, inc):
    """BLAS element index (src/mpvector.cuh:65-67)"""
    return k * inc if inc > 0 else (-n + k + 1) * inc


def val(rec, m):
    return exact.mp_value(rec, m)


def assert_enclosed(rec, m):
    X = exact.crt(rec["digits"], m)
    q = Fraction(X, exact.moduli_product(m))
    assert exact.er_value(rec["eval"][0]) <= q <= exact.er_value(rec["eval"][1])


def mods(o):
    return [int(v) for v in o.moduli()]


# ---------------------------------------------------------------- exact values
@pytest.fixture(scope="module")
def o32():
    return po.Oracle("32_8double")


@pytest.fixture(scope="module")
def o16():
    return po.Oracle("16_4double")


def test_dot_exact(o32):
    m = mods(o32)
    n = 300
    x = o32.set_d(gen.uexp(1, n, -10, 10))
    y = o32.set_d(gen.uexp(2, n, -10, 10))
    X = [val(v, m) for v in x]
    Y = [val(v, m) for v in y]
    for g3, b3 in ((1, 1), (4, 8), (7, 33)):
        r = o32.dot(x, y, g3, b3)
        assert val(r[0], m) == sum(a * b for a, b in zip(X, Y))
        assert_enclosed(r[0], m)
    # negative and non-unit strides
    nn = 100
    r = o32.dot(x, y, 4, 8, incx=-2, incy=3, n=nn)
    assert val(r[0], m) == sum(X[bidx(i, nn, -2)] * Y[bidx(i, nn, 3)] for i in range(nn))


def test_axpy_exact(o32):
    m = mods(o32)
    n = 120
    x = o32.set_d(gen.uexp(3, n, -10, 10))
    y = o32.set_d(gen.uexp(4, n, -10, 10))
    alpha = o32.set_d([-2.5])
    a = val(alpha[0], m)
    X = [val(v, m) for v in x]
    Y = [val(v, m) for v in y]
    r = o32.axpy(alpha, x, y)
    assert [val(v, m) for v in r] == [a * xi + yi for xi, yi in zip(X, Y)]
    r2 = o32.axpy_v2(alpha, x, y)
    assert [val(v, m) for v in r2] == [a * xi + yi for xi, yi in zip(X, Y)]
    nn = 50
    r = o32.axpy(alpha, x, y, incx=2, incy=-1, n=nn)
    want = list(Y)
    for i in range(nn):
        want[bidx(i, nn, -1)] = a * X[bidx(i, nn, 2)] + Y[bidx(i, nn, -1)]
    assert [val(v, m) for v in r] == want


@pytest.mark.parametrize("trans", [0, 1])
def test_gemv_exact(o32, trans):
    m_ = mods(o32)
    m, n, lda = 7, 9, 10  # column-major A with padding rows
    A = o32.set_d(gen.uexp(5, lda * n, -10, 10))
    lx, ly = (m, n) if trans else (n, m)
    x = o32.set_d(gen.uexp(6, lx, -10, 10))
    y = o32.set_d(gen.uexp(7, ly, -10, 10))
    alpha, beta = o32.set_d([0.75]), o32.set_d([-3.0])
    Av = [val(v, m_) for v in A]
    xv = [val(v, m_) for v in x]
    yv = [val(v, m_) for v in y]
    a, b = val(alpha[0], m_), val(beta[0], m_)
    r = o32.gemv(trans, m, n, alpha, A, lda, x, beta, y, 4)
    for o in range(ly):
        if trans:  # y_o = sum_i A[i, o] x_i
            s = sum(Av[o * lda + i] * xv[i] for i in range(m))
        else:
            s = sum(Av[j * lda + o] * xv[j] for j in range(n))
        assert val(r[o], m_) == a * s + b * yv[o], o
        assert_enclosed(r[o], m_)


def test_gemm_exact(o32):
    m_ = mods(o32)
    m, n, k = 4, 3, 5
    lda, ldb, ldc = 6, 7, 5
    A = o32.set_d(gen.uexp(8, lda * k, -10, 10))
    B = o32.set_d(gen.uexp(9, ldb * n, -10, 10))
    Cm = o32.set_d(gen.uexp(10, ldc * n, -10, 10))
    alpha, beta = o32.set_d([1.25]), o32.set_d([-0.5])
    Av, Bv, Cv = [val(v, m_) for v in A], [val(v, m_) for v in B], [val(v, m_) for v in Cm]
    a, b = val(alpha[0], m_), val(beta[0], m_)
    r = o32.gemm(m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc)
    for j in range(n):
        for i in range(m):
            s = sum(Av[p * lda + i] * Bv[j * ldb + p] for p in range(k))
            assert val(r[j * ldc + i], m_) == a * s + b * Cv[j * ldc + i]
    # v2 with every transpose combination: op(A) is m x k, op(B) is k x n
    for ta in (0, 1):
        for tb in (0, 1):
            la = k if ta else m
            lb = n if tb else k
            r = o32.gemm_v2(ta, tb, m, n, k, alpha, A, la, B, lb, beta, Cm, m)
            for j in range(n):
                for i in range(m):
                    s = 0
                    for p in range(k):
                        av = Av[p + la * i] if ta else Av[i + la * p]
                        bv = Bv[j + lb * p] if tb else Bv[p + lb * j]
                        s += av * bv
                    assert val(r[i + m * j], m_) == a * s + b * Cv[i + m * j], (ta, tb, i, j)


@pytest.mark.parametrize("tpr", [1, 2, 8, 32])
def test_spmv_exact(o16, tpr):
    m_ = mods(o16)
    m = 60
    irp, ja, as_ = gen.powerlaw_csr(11, m, kmax=40)
    x = o16.set_d(gen.uexp(12, m, -10, 10))
    xv = [val(v, m_) for v in x]
    y = o16.spmv(m, irp, ja, as_, x, tpr)
    for r in range(m):
        s = sum(Fraction(float(as_[i])) * xv[ja[i]] for i in range(irp[r], irp[r + 1]))
        assert val(y[r], m_) == s, r


def test_vadd_closed_form(o32):
    """v1 vector add (src/mpvector.cuh:234-472): exact aligned sum when the
    interval sum does not straddle zero and nothing rounds."""
    m_ = mods(o32)
    x = o32.set_d(gen.uexp(13, 400, -40, 40))
    y = o32.set_d(gen.uexp(14, 400, -40, 40))
    r = o32.vadd(x, y)
    for a, b, c in zip(x, y, r):
        assert val(c, m_) == val(a, m_) + val(b, m_)
        assert_enclosed(c, m_)


# ---------------------------------------------------------------- summation order
@pytest.fixture(scope="module")
def o8():
    return po.Oracle("8_2double")


def one(o, rec):
    a = o.empty(1)
    a[0] = rec
    return a


def zeros(o, count):
    return o.set_d(np.zeros(count))


def halving_tree(o, part, b):
    """src/mpreduct.cuh:60-68: for i = nextPow2(b)/2 .. 1: s[t] += s[t+i] (t < i, t+i < b)"""
    part = part.copy()
    p2 = 1
    while p2 < b:
        p2 *= 2
    i = p2 // 2
    while i >= 1:
        t = np.arange(i)
        t = t[t + i < b]
        if len(t):
            part[t] = o.add(part[t], part[t + i])
        i //= 2
    return part[0]


def strided_fold(o, terms, nvt):
    """thread v accumulates terms v, v + nvt, ... from zero (src/mpreduct.cuh:38-56)"""
    acc = zeros(o, nvt)
    for s in range(0, len(terms), nvt):
        cnt = min(nvt, len(terms) - s)
        acc[:cnt] = o.add(acc[:cnt], terms[s:s + cnt])
    return acc


def test_dot_order(o8):
    n, g3, b3 = 1000, 5, 24
    x = o8.set_d(gen.uexp(21, n, -20, 20))
    y = o8.set_d(gen.uexp(22, n, -20, 20))
    prods = o8.mul(x, y)  # mp_vec2vec_mul, src/blas/dot.cuh:99
    part = strided_fold(o8, prods, g3 * b3)
    blocks = np.concatenate([one(o8, halving_tree(o8, part[g * b3:(g + 1) * b3], b3)) for g in range(g3)])
    # second pass: one block of b3 threads over the g3 block results (src/mpreduct.cuh:81-112)
    part2 = strided_fold(o8, blocks, b3)
    want = halving_tree(o8, part2, b3)
    got = o8.dot(x, y, g3, b3)[0]
    assert got.tobytes() == want.tobytes()


def test_gemv_order(o8):
    m, n, b3 = 6, 50, 8
    A = o8.set_d(gen.uexp(23, m * n, -20, 20))
    x = o8.set_d(gen.uexp(24, n, -20, 20))
    y = o8.set_d(gen.uexp(25, m, -20, 20))
    alpha, beta = o8.set_d([1.0 / 3.0]), o8.set_d([-7.0])
    bx = o8.mul(x, np.repeat(alpha, n))  # src/blas/gemv.cuh:175-181
    by = o8.mul(y, np.repeat(beta, m))   # :184-190
    want = []
    for o in range(m):
        prods = o8.mul(A[o::m][:n], bx)  # row o of column-major A times buffer1
        s = halving_tree(o8, strided_fold(o8, prods, b3), b3)  # :44-76
        want.append(o8.add(one(o8, by[o]), one(o8, s))[0])
    got = o8.gemv(0, m, n, alpha, A, m, x, beta, y, b3)
    assert got.tobytes() == np.array(want, dtype=got.dtype).tobytes()


def test_gemm_order(o8):
    m, n, k = 3, 2, 30
    A = o8.set_d(gen.uexp(26, m * k, -20, 20))
    B = o8.set_d(gen.uexp(27, k * n, -20, 20))
    Cm = o8.set_d(gen.uexp(28, m * n, -20, 20))
    alpha, beta = o8.set_d([0.1]), o8.set_d([2.0 / 3.0])
    want1, want2 = Cm.copy(), Cm.copy()
    for j in range(n):
        for i in range(m):
            s = zeros(o8, 1)
            for p in range(k):  # serial k order, src/blas/gemm.cuh:39-58
                s = o8.add(s, o8.mul(one(o8, A[i + m * p]), one(o8, B[p + k * j])))
            buf = o8.mul(s, alpha)
            cb = o8.mul(one(o8, Cm[i + m * j]), beta)
            want1[i + m * j] = o8.vadd(cb, buf)[0]  # :127-166
            want2[i + m * j] = o8.add(buf, cb)[0]  # src/blas/v2/gemm_v2.cuh:76-78
    got1 = o8.gemm(m, n, k, alpha, A, m, B, k, beta, Cm, m)
    got2 = o8.gemm_v2(0, 0, m, n, k, alpha, A, m, B, k, beta, Cm, m)
    assert got1.tobytes() == want1.tobytes()
    assert got2.tobytes() == want2.tobytes()


@pytest.mark.parametrize("tpr", [1, 4, 32])
def test_spmv_order(o8, tpr):
    m = 40
    irp, ja, as_ = gen.powerlaw_csr(29, m, kmax=64)
    x = o8.set_d(gen.uexp(30, m, -20, 20))
    got = o8.spmv(m, irp, ja, as_, x, tpr)
    for r in range(m):
        lanes = []
        for lane in range(tpr):  # lane-strided partial sums, src/sparse/spmv/spmv_csrv.cuh:61-65
            acc = zeros(o8, 1)
            for i in range(irp[r] + lane, irp[r + 1], tpr):
                acc = o8.add(acc, o8.mul_d(one(o8, x[ja[i]]), [as_[i]]))
            lanes.append(acc)
        h = tpr // 2
        while h >= 1:  # 16/8/4/2/1 tree, :66-80
            for t in range(h):
                lanes[t] = o8.add(lanes[t], lanes[t + h])
            h //= 2
        assert got[r].tobytes() == lanes[0][0].tobytes(), r


def test_dot_rounds_within_bound(o8):
    """At 106 bits the dot result is within n roundings of the exact value."""
    m_ = mods(o8)
    n = 2000
    x = o8.set_d(gen.uexp(31, n, -20, 20))
    y = o8.set_d(gen.uexp(32, n, -20, 20))
    r = o8.dot(x, y, 8, 16)[0]
    terms = [val(a, m_) * val(b, m_) for a, b in zip(x, y)]
    ex = sum(terms)
    bound = sum(abs(t) for t in terms) * 2 * n * Fraction(2) ** (-o8.mp_consts()[0])
    assert abs(val(r, m_) - ex) <= bound
