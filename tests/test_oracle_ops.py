"""Pins the oracle's scalar algorithms against what the mathematics fixes:
exact rational rounding, CRT closed forms, enclosure invariants, and the
reference's own verify operands."""
import math
import struct
from fractions import Fraction

import numpy as np
import pytest

import exact
import pyoracle as po

D, E = po.POLICY_D, po.POLICY_E


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def rand_doubles(rng, count, lo=-60, hi=60):
    u = rng.uniform(-1, 1, count)
    e = rng.integers(lo, hi, count)
    return np.ldexp(u, e)


# ---------------------------------------------------------------- rounding
@pytest.mark.parametrize("op", ["add", "sub", "mul", "div"])
def test_policy_d_directed_ops_exact(oracle_lib, op):
    L = oracle_lib.lib()
    fn = {"add": L.orc_dadd, "sub": L.orc_dsub, "mul": L.orc_dmul, "div": L.orc_ddiv}[op]
    rng = np.random.default_rng(1)
    a = rand_doubles(rng, 3000)
    b = rand_doubles(rng, 3000)
    b[:50] = a[:50]  # exact cancellations / exact quotients
    for x, y in zip(a, b):
        x, y = float(x), float(y)
        q = {"add": Fraction(x) + Fraction(y), "sub": Fraction(x) - Fraction(y), "mul": Fraction(x) * Fraction(y),
             "div": Fraction(x) / Fraction(y)}[op]
        for up in (0, 1):
            got = fn(D, up, x, y)
            want = exact.round_dir(q, bool(up))
            if q == 0:
                # IEEE: an exact zero sum is -0 under round-down, +0 otherwise (x - x, x + (-x))
                if op in ("add", "sub"):
                    want = -0.0 if not up else 0.0
            assert bits(got) == bits(want), (op, x, y, up, got, want)


def test_policy_e_directed_ops_enclose(oracle_lib):
    L = oracle_lib.lib()
    rng = np.random.default_rng(2)
    a = rand_doubles(rng, 2000)
    b = rand_doubles(rng, 2000)
    for x, y in zip(a, b):
        x, y = float(x), float(y)
        for fn, q in ((L.orc_dadd, Fraction(x) + Fraction(y)), (L.orc_dmul, Fraction(x) * Fraction(y)),
                      (L.orc_ddiv, Fraction(x) / Fraction(y))):
            lo, hi = fn(E, 0, x, y), fn(E, 1, x, y)
            assert Fraction(lo) < q < Fraction(hi)
            rn = float(q)
            assert math.nextafter(math.nextafter(rn, -math.inf), -math.inf) <= lo
            assert hi <= math.nextafter(math.nextafter(rn, math.inf), math.inf)


@pytest.mark.parametrize("n", [8, 16, 32])
def test_psum_tree_shape(oracle_lib, n):
    """psum_rd/ru: balanced pairwise tree over n values plus a root '+0'
    (src/pairwise.cuh psum(2n)<n>), each node directed-rounded."""
    L = oracle_lib.lib()
    rng = np.random.default_rng(n)
    for _ in range(200):
        x = np.abs(rng.uniform(0, 1, n)) * rng.choice([1e-8, 1.0, 1e8], n)

        def tree(v, up):
            if len(v) == 1:
                return v[0]
            h = len(v) // 2
            return exact.round_dir(Fraction(tree(v[:h], up)) + Fraction(tree(v[h:], up)), up)

        for up in (0, 1):
            want = tree([float(t) for t in x], bool(up))
            got = L.orc_psum(D, up, n, x.ctypes.data_as(po.C.POINTER(po.C.c_double)))
            assert bits(got) == bits(want)
            e = L.orc_psum(E, up, n, x.ctypes.data_as(po.C.POINTER(po.C.c_double)))
            exact_sum = sum(Fraction(float(t)) for t in x)
            assert (Fraction(e) >= exact_sum) if up else (Fraction(e) <= exact_sum)


# ---------------------------------------------------------------- RNS integer algorithms
@pytest.fixture(scope="module", params=["8_2double", "16_4double", "32_8double", "8_32bit"])
def orc(request, oracle_lib):
    return po.Oracle(request.param)


def rand_X(rng, M, count):
    """Random integers in [0, M): uniform, tiny, and close to M."""
    out = []
    for i in range(count):
        kind = i % 4
        if kind == 0:
            out.append(int.from_bytes(rng.bytes(M.bit_length() // 8 + 8), "little") % M)
        elif kind == 1:
            out.append(int(rng.integers(0, 2**62)) >> int(rng.integers(0, 60)))
        elif kind == 2:
            out.append(M - 1 - int(rng.integers(0, 2**40)))
        else:
            out.append(int.from_bytes(rng.bytes(16), "little") << int(rng.integers(0, M.bit_length() - 130)))
    return [x % M for x in out]


def test_mrd_and_rank(orc):
    m = [int(v) for v in orc.moduli()]
    M = exact.moduli_product(m)
    w = [int(v) for v in orc.part_inverse()]
    rng = np.random.default_rng(5)
    Xs = rand_X(rng, M, 200)
    dig = np.array([exact.residues(X, m) for X in Xs], np.int32)
    mrd = orc.mrd(dig)
    for X, r in zip(Xs, mrd):
        assert r == exact.mrc_digits(X, m)[-1] == X // (M // m[-1])
    for pol in (D, E):
        k = orc.rank(dig, fast=False, policy=pol)
        for X, kk, d in zip(Xs, k, dig):
            assert kk == exact.exact_rank(d, m, w)
    kf = orc.rank(dig, fast=True, policy=D)
    for X, kk, d in zip(Xs, kf, dig):
        if X < M // 2:
            assert kk == exact.exact_rank(d, m, w)


def test_scale2pow_closed_form(orc):
    m = [int(v) for v in orc.moduli()]
    M = exact.moduli_product(m)
    rng = np.random.default_rng(6)
    Xs = rand_X(rng, M, 240)
    Ds = [int(rng.integers(1, M.bit_length())) for _ in Xs]
    dig = np.array([exact.residues(X, m) for X in Xs], np.int32)
    for pol in (D, E):
        r = orc.scale2pow(dig, np.array(Ds, np.uint32), policy=pol)
        for X, Dd, rr in zip(Xs, Ds, r):
            if pol == D and Dd < 30 and X > M // 2:
                continue  # the CUDA single-step path uses the fast rank, exact only for X/M < 1/2
            if pol == E:
                # the CPU path's double-reciprocal products may leave residues non-canonical
                rr = [int(v) % mi for v, mi in zip(rr, m)]
            assert exact.crt(rr, m) == X >> Dd, (X, Dd)


def test_eval_encloses(orc):
    m = [int(v) for v in orc.moduli()]
    M = exact.moduli_product(m)
    rng = np.random.default_rng(7)
    Xs = rand_X(rng, M, 200) + [1, 2, M - 1, M - 2, M // 2]
    dig = np.array([exact.residues(X, m) for X in Xs], np.int32)
    acc = orc.eval_consts()["acc"]
    small = [i for i, X in enumerate(Xs) if X <= M // 2]
    for pol in (D, E):
        for fast in (False, True):
            # the fast evaluation is only defined for numbers that are not too large
            # (its refinement loop need not terminate for X close to M)
            sel = small if fast else list(range(len(Xs)))
            lo, up = orc.eval(dig[sel], fast=fast, policy=pol)
            for X, l, u in zip([Xs[i] for i in sel], lo, up):
                q = Fraction(X, M)
                lv, uv = exact.er_value(l), exact.er_value(u)
                assert lv <= q <= uv, (pol, fast, X)
                if X > 0 and not (q > Fraction(1, 2)):
                    # refinement: relative width bounded by the accuracy target
                    assert (uv - lv) / q < 4 * Fraction(acc) if q < Fraction(1, 4) else True
    zl, zu = orc.eval(np.zeros((1, len(m)), np.int32), policy=D)
    assert zl[0]["frac"] == 0 and zu[0]["frac"] == 0 and zl[0]["exp"] == 0


def test_eval_walk_down_sweep(orc):
    """tests/arith/test_rns_eval_verify.cu:77-131 walks X down from M-1 and compares
    the evaluation with the exact fraction; here as an asserted property."""
    m = [int(v) for v in orc.moduli()]
    M = exact.moduli_product(m)
    Xs = [M - 1 - i for i in range(0, 64)] + [M // (2**k) for k in range(1, 200, 7)]
    dig = np.array([exact.residues(X, m) for X in Xs], np.int32)
    for pol in (D, E):
        lo, up = orc.eval(dig, policy=pol)
        for X, l, u in zip(Xs, lo, up):
            assert exact.er_value(l) <= Fraction(X, M) <= exact.er_value(u)


def test_rns_cmp(orc):
    m = [int(v) for v in orc.moduli()]
    M = exact.moduli_product(m)
    rng = np.random.default_rng(8)
    Xs = rand_X(rng, M, 60)
    Ys = [X if i % 5 == 0 else (X + 1) % M if i % 5 == 1 else Y for i, (X, Y) in enumerate(zip(Xs, rand_X(rng, M, 60)))]
    dx = np.array([exact.residues(X, m) for X in Xs], np.int32)
    dy = np.array([exact.residues(Y, m) for Y in Ys], np.int32)
    for pol in (D, E):
        c = orc.rns_cmp(dx, dy, policy=pol)
        for X, Y, cc in zip(Xs, Ys, c):
            assert cc == (X > Y) - (X < Y)


# ---------------------------------------------------------------- mp_float_t ops
def test_set_d_closed_form(orc):
    m = [int(v) for v in orc.moduli()]
    rng = np.random.default_rng(9)
    x = np.concatenate([rand_doubles(rng, 300, -300, 300), [0.0, -0.0, 1.0, -1.0, 0.5, 3.0, 2.0**-60, 1e300]])
    for pol in (D, E):
        r = orc.set_d(x, policy=pol)
        for xv, rec in zip(x, r):
            assert exact.mp_value(rec, m) == Fraction(float(xv))
            X = exact.crt(rec["digits"], m)
            assert X == 0 or X % 2 == 1  # trailing zeros trimmed
            assert rec["sign"] == (math.copysign(1, xv) < 0)
            lv, uv = exact.er_value(rec["eval"][0]), exact.er_value(rec["eval"][1])
            q = Fraction(X, exact.moduli_product(m))
            assert lv <= q <= uv


def test_reference_verify_operands(oracle_lib):
    """tests/arith/test_mp_add_verify.cu:47-48 and test_mp_mul_verify.cu:42-43."""
    for name in ("8_2double", "16_4double", "32_8double"):
        o = po.Oracle(name)
        m = [int(v) for v in o.moduli()]
        for pol in (D, E):
            x = o.set_d([-1000.01], policy=pol)
            y = o.set_d([10000.000000002], policy=pol)
            s = o.add(x, y, policy=pol)
            assert exact.mp_value(s[0], m) == Fraction(-1000.01) + Fraction(10000.000000002)
            d = o.sub(x, y, policy=pol)
            assert exact.mp_value(d[0], m) == Fraction(-1000.01) - Fraction(10000.000000002)
            one = o.set_d([1.0], policy=pol)
            mone = o.set_d([-1.0], policy=pol)
            p = o.mul(one, mone, policy=pol)
            assert exact.mp_value(p[0], m) == -1
            assert list(p[0]["digits"]) == [1] * len(m) and p[0]["sign"] == 1 and p[0]["exp"] == 0
            q = o.mul_d(one, [-1.0], policy=pol)
            assert exact.mp_value(q[0], m) == -1


def _aligned_add_closed_form(x, y, m, mp_j):
    """Value of mp_add given the operand-zeroing decisions of src/arith/add.cuh:37-56."""
    ex, ey = int(x["exp"]), int(y["exp"])
    dexp = ex - ey
    gamma, theta = max(dexp, 0), max(-dexp, 0)
    nzx = float(y["eval"][1]["frac"]) == 0 or theta + int(y["eval"][1]["exp"]) < mp_j
    nzy = float(x["eval"][1]["frac"]) == 0 or gamma + int(x["eval"][1]["exp"]) < mp_j
    vx = exact.mp_value(x, m) if nzx else 0
    vy = exact.mp_value(y, m) if nzy else 0
    return vx + vy


@pytest.mark.parametrize("pol", [D, E])
def test_add_mul_closed_form(orc, pol):
    m = [int(v) for v in orc.moduli()]
    prec, mp_h, mp_j = orc.mp_consts()
    rng = np.random.default_rng(10 + pol)
    cnt = 300
    a = orc.set_d(rand_doubles(rng, cnt, -40, 40), policy=pol)
    b = orc.set_d(rand_doubles(rng, cnt, -40, 40), policy=pol)
    # products (106-bit significands) exercise rounding at n = 8
    pa = orc.mul(a, b, policy=pol)
    pb = orc.mul(b[::-1].copy(), a, policy=pol)
    for x, y, r in zip(a, b, pa):
        want = exact.mp_value(x, m) * exact.mp_value(y, m)
        got = exact.mp_value(r, m)
        if int(r["exp"]) == int(x["exp"]) + int(y["exp"]):
            assert got == want
        else:  # rounded: truncation of |X*Y| by (r.exp - x.exp - y.exp) bits
            sh = int(r["exp"]) - int(x["exp"]) - int(y["exp"])
            assert sh > 0 and abs(got) == Fraction((abs(want) / Fraction(2) ** (int(x["exp"]) + int(y["exp"]))).numerator >> sh) * Fraction(2) ** int(r["exp"])
    s = orc.add(pa, pb, policy=pol)
    for x, y, r in zip(pa, pb, s):
        want = _aligned_add_closed_form(x, y, m, mp_j)
        got = exact.mp_value(r, m)
        emin = min(int(x["exp"]), int(y["exp"]))
        if want == 0:
            assert got == 0
            continue
        scale = Fraction(2) ** int(r["exp"])
        # result = truncation of the exact aligned sum at the result exponent
        assert int(r["sign"]) == (want < 0)
        assert abs(got) == Fraction(int(abs(want) / scale)) * scale
        X = exact.crt(r["digits"], m)
        lv, uv = exact.er_value(r["eval"][0]), exact.er_value(r["eval"][1])
        q = Fraction(X, exact.moduli_product(m))
        assert lv <= q <= uv


@pytest.mark.parametrize("pol", [D, E])
def test_mp_cmp_exact(orc, pol):
    """mp_cmp / mp_cmp_abs (src/arith/cmp.cuh, cmpabs.cuh) against the exact comparison of the
    values, on separated, equal, near-equal (mixed-radix decision), dropped and zero operands.
    Two zeros compare by sign in the reference (+0 > -0: cmp.cuh:64-65), so they are checked
    against that rule."""
    m = [int(v) for v in orc.moduli()]
    rng = np.random.default_rng(31 + pol)
    cnt = 200
    a = orc.set_d(rand_doubles(rng, cnt, -300, 300), policy=pol)
    b = orc.set_d(rand_doubles(rng, cnt, -300, 300), policy=pol)
    p = orc.mul(a, b, policy=pol)
    neg = a.copy()
    neg["sign"] ^= 1
    z = orc.set_d(np.array([0.0, -0.0] * (cnt // 2)), policy=pol)
    near = orc.add(p, orc.mul(p, orc.set_d(np.full(cnt, 2.0**-80), policy=pol), policy=pol), policy=pol)
    # the same value with another exponent (2^40 through the residues and back)
    same = orc.mul(orc.mul(a, orc.set_d(np.full(cnt, 2.0**40), policy=pol), policy=pol),
                   orc.set_d(np.full(cnt, 2.0**-40), policy=pol), policy=pol)
    X = np.concatenate([a, a, a, p, near, z, z, same, a])
    Y = np.concatenate([b, a, neg, near, p, z, z[::-1], a, z])
    c = orc.cmp(X, Y, False, pol)
    ca = orc.cmp(X, Y, True, pol)
    seen = set()
    for x, y, cc, cca in zip(X, Y, c, ca):
        vx, vy = exact.mp_value(x, m), exact.mp_value(y, m)
        if vx == 0 and vy == 0:
            want = int(y["sign"]) - int(x["sign"])
        else:
            want = (vx > vy) - (vx < vy)
        assert cc == want, (float(vx), float(vy), cc)
        assert cca == (abs(vx) > abs(vy)) - (abs(vx) < abs(vy))
        seen.add(cc)
    assert seen == {-1, 0, 1}
