"""The oracle's solver-path functions pinned without the restatement itself: mp_div,
mp_sqrt and mp_dsqrt against exact integer arithmetic (the MPFR results they stand for),
mp_add_abs against its closed form, the norm2 sum-of-squares order against the v2 dot
of x with itself, and the composed CG / PCG against the exact residual of its own
iterate."""
import math
from fractions import Fraction
import os

import numpy as np
import pytest

import exact
import gen
import pyoracle as po

D, E = po.POLICY_D, po.POLICY_E
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SETS = ["8_2double", "16_4double", "32_8double"]


def mods(o):
    return [int(v) for v in o.moduli()]


def round_int(X, p):
    """X (> 0) rounded to p significant bits to nearest, ties to even: (Q, shift)"""
    b = X.bit_length()
    if b <= p:
        return X, 0
    q, r = divmod(X, 1 << (b - p))
    h = 1 << (b - p - 1)
    if r > h or (r == h and q & 1):
        q += 1
    return q, b - p


def get_mpfr(rec, m, p):
    """(-1)^s X 2^e rounded to p bits: (sign, Q, e)"""
    X = exact.crt(rec["digits"], m)
    q, sh = round_int(X, p)
    return int(rec["sign"]), q, int(rec["exp"]) + sh


def rational_round(v, p):
    """a positive rational rounded to p significant bits to nearest (ties to even): Q 2^e"""
    e = v.numerator.bit_length() - v.denominator.bit_length() - p - 2
    while Fraction(2) ** (e + p + 2) <= v:
        e += 1
    while Fraction(2) ** (e + p + 1) > v:
        e -= 1
    s = v / Fraction(2) ** e  # in [2^(p+1), 2^(p+2))
    q = s.numerator // s.denominator
    rem = s - q
    # q has p + 2 bits; drop 2 with the remainder as sticky
    low = q & 3
    q >>= 2
    if low > 2 or (low == 2 and (rem != 0 or q & 1)):
        q += 1
    return q, e + 2


def set_mpfr_fields(sign, q, e):
    """mp_set_mpfr's sign, exponent and significand (trailing zeros dropped) of (-1)^sign q 2^e"""
    if q == 0:
        return sign, (0 if sign else -1), 0
    tz = (q & -q).bit_length() - 1
    return sign, e + tz, q >> tz


def mp_precision(name):
    return {"8_2double": 106, "16_4double": 212, "32_8double": 424}[name]


def operands(o, seed, count):
    """full-length significands: products and sums of converted doubles, plus plain doubles"""
    a = o.set_d(gen.uexp(seed, count, -40, 40), D)
    b = o.set_d(gen.uexp(seed + 1, count, -40, 40), D)
    return np.concatenate([o.mul(a, b, D), o.add(a, b, D), a])


@pytest.mark.parametrize("name", SETS)
def test_div_against_exact(name):
    o = po.Oracle(name)
    m, P = mods(o), mp_precision(name)
    x, y = operands(o, 11, 60), operands(o, 21, 60)[::-1].copy()
    x[:3] = o.set_d([0.0, -0.0, 5.0], D)
    y[2] = x[2]
    r = o.div(x, y)
    for i in range(len(x)):
        sx, qx, ex = get_mpfr(x[i], m, P)
        sy, qy, ey = get_mpfr(y[i], m, P)
        s = sx ^ sy
        if qx == 0:
            want = set_mpfr_fields(s, 0, 0)
        else:
            q, e = rational_round(Fraction(qx, qy) * Fraction(2) ** (ex - ey), P)
            want = set_mpfr_fields(s, q, e)
        got = (int(r[i]["sign"]), int(r[i]["exp"]), exact.crt(r[i]["digits"], m))
        assert got == want, f"{name} record {i}"
        # the evaluation is the CPU path's: it encloses X / M
        if want[2]:
            Q = Fraction(want[2], exact.moduli_product(m))
            assert exact.er_value(r[i]["eval"][0]) <= Q <= exact.er_value(r[i]["eval"][1])


@pytest.mark.parametrize("name", SETS)
def test_sqrt_and_dsqrt_against_exact(name):
    o = po.Oracle(name)
    m, P = mods(o), mp_precision(name)
    x = operands(o, 31, 60)
    x["sign"] = 0
    r, d = o.sqrt(x), o.dsqrt(x)
    for i in range(len(x)):
        _, qx, ex = get_mpfr(x[i], m, P)
        if qx == 0:
            assert d[i] == 0.0
            continue
        # sqrt(qx 2^ex): an even exponent, the radicand widened, integer square root
        sh = max(0, 2 * P + 2 - qx.bit_length())
        if (ex - sh) & 1:
            sh += 1
        rad = qx << sh
        s = math.isqrt(rad)
        es = (ex - sh) // 2
        inexact = s * s != rad
        extra = s.bit_length() - P  # 1 or 2 bits dropped
        low, s = s & ((1 << extra) - 1), s >> extra
        h = 1 << (extra - 1)
        if low > h or (low == h and (inexact or s & 1)):
            s += 1
        e = es + extra
        want = set_mpfr_fields(0, s, e)
        got = (int(r[i]["sign"]), int(r[i]["exp"]), exact.crt(r[i]["digits"], m))
        assert got == want, f"{name} record {i}"
        assert d[i] == float(Fraction(s) * Fraction(2) ** e)


@pytest.mark.parametrize("name", SETS)
@pytest.mark.parametrize("policy", [D, E])
def test_add_abs_closed_form(name, policy):
    """|x| + |y| on the exact values: the residues hold |x| 2^(ex - e) + |y| 2^(ey - e) (or the
    rounded sum) and the bounds enclose it"""
    o = po.Oracle(name)
    m = mods(o)
    x, y = operands(o, 41, 40), operands(o, 51, 40)
    r = o.add_abs(x, y, policy)
    for i in range(len(x)):
        want = abs(exact.mp_value(x[i], m)) + abs(exact.mp_value(y[i], m))
        got = exact.mp_value(r[i], m)
        assert r[i]["sign"] == 0
        if got != want:  # rounded: within the precision
            assert abs(got - want) <= want * Fraction(2) ** (-mp_precision(name) + 2)


@pytest.mark.parametrize("name", SETS)
@pytest.mark.parametrize("blocks,threads", [(256, 32), (7, 16), (3, 5)])
def test_sumsq_is_the_dot_of_x_with_itself(name, blocks, threads):
    """mp_norm2's sum of squares (mp_add_abs) equals the v2 dot of x with itself (mp_add) bit
    for bit: the squares and the zero start are nonnegative, and on nonnegative operands
    mp_add_abs and mp_add take the same steps"""
    o = po.Oracle(name)
    x = o.set_d(gen.uexp(61, 3000, -60, 60), D)
    x[::7] = o.set_d(np.zeros(len(x[::7])), D)
    for pol in (D, E):
        a = o.sumsq(x, blocks, threads, pol)
        b = o.dot(x, x, blocks, threads, pol)
        assert a.tobytes() == b.tobytes()


def _lf10():
    return gen.read_mtx(os.path.join(ROOT, "tests", "golden", "LF10.mtx"))


@pytest.mark.parametrize("name", ["8_2double", "32_8double"])
@pytest.mark.parametrize("precond", [False, True])
def test_cg_solves_lf10(name, precond):
    """the composed CG / PCG on the reference's LF10 matrix: the relative residual history
    ends below tol, and the exact residual of the returned iterate is small"""
    o = po.Oracle(name)
    m = mods(o)
    n, _, irp, ja, as_ = _lf10()
    xt = gen.uexp(71, n, -2, 2)
    b = o.spmv(n, irp, ja, as_, o.set_d(xt, D), 1, D)
    x0 = o.set_d(np.zeros(n), D)
    diag = np.array([as_[irp[i] + list(ja[irp[i]:irp[i + 1]]).index(i)] for i in range(n)])
    M = 1.0 / diag if precond else None
    tol = 1e-20 if name == "8_2double" else 1e-60
    x, res = o.cg_csr(n, irp, ja, as_, b, tol, 500, x0, precond=M)
    assert res[0] == 1.0 and res[-1] <= tol
    # exact residual ||b - A x|| / ||b||
    xv = [exact.mp_value(v, m) for v in x]
    bv = [exact.mp_value(v, m) for v in b]
    rr = 0
    for i in range(n):
        s = sum(Fraction(float(as_[q])) * xv[ja[q]] for q in range(irp[i], irp[i + 1]))
        rr += (bv[i] - s) ** 2
    nb = sum(v * v for v in bv)
    assert float(rr / nb) < (1e3 * tol) ** 2
