"""Bits of the oracle's directed operations and of its interval evaluation, pinned without the
oracle's own code:

* policy E (the reference's CPU path): every directed add, sub, mul and div equals its
  definition RN(a op b) -/+ (phi |c| + eta), with phi = DBL_EPS (1 + 2 DBL_EPS) and
  eta = DBL_ETA derived here from src/bitwise.cuh:51-61 and the unfused fmac of
  src/dinterval.cuh:40-48 (EMPLOY_STD_FMA false in every params file), on edge inputs;
  the reciprocal tables of rns_const_init (src/rns.cuh:413-414) likewise;
* the evaluation (src/rns.cuh:645-781, CUDA :797-933) recomposed in Python from exact
  rational directed rounding (policy D), the definition above (policy E), the moduli and
  the CRT, bit for bit, on inputs that force every branch: zero, the mr > 0 and mr == 0
  ambiguity branches, one-step and multi-step refinement (the REF_FACTOR clamp cannot fire:
  shown per set below);
* fast == full wherever no ambiguity branch fires, and the width of the bounds within the
  rounding-error bound of the sums (N (3 + log2 N) 2^-52 absolute, twice that for the
  emulated rounding) where no refinement runs;
* the variant each operation uses (full in mp_set_d, fast in mp_round) by the eval bits of
  its results."""
import math
import random
import struct
from fractions import Fraction

import numpy as np
import pytest

import exact
import gen
import pyoracle as po

D, E = po.POLICY_D, po.POLICY_E
EPS = math.ldexp(1.0, -53)  # DBL_EPS
ETA = math.ldexp(1.0, -1074)  # DBL_ETA
PHI1 = EPS * (1 + 2 * EPS)  # DBL_PHI1
SETS = ["8_2double", "16_4double", "32_8double", "8_32bit"]


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


def e_op(op, a, b, up):
    """the CPU path's directed operation, by its definition (Python floats are binary64, RN)"""
    c = a + b if op == "add" else a - b if op == "sub" else a * b if op == "mul" else a / b
    e = PHI1 * abs(c) + ETA
    return c + e if up else c - e


def edge_doubles(rng, count):
    out = [0.0, -0.0, ETA, -ETA, 2 * ETA, math.ldexp(1, -1022), math.ldexp(1, -1023), 1.0, -1.0, 3.0, 0.1,
           math.ldexp(1, 1000), -math.ldexp(1, 1000), 1.7976931348623157e308 / 4, math.nextafter(1.0, 2)]
    u = rng.uniform(-1, 1, count)
    e = rng.integers(-1074, 1000, count)
    return out + list(np.ldexp(u, e)) + list(np.ldexp(u, rng.integers(-60, 60, count)))


@pytest.mark.parametrize("op", ["add", "sub", "mul", "div"])
def test_policy_e_ops_are_their_definition(oracle_lib, op):
    L = oracle_lib.lib()
    fn = {"add": L.orc_dadd, "sub": L.orc_dsub, "mul": L.orc_dmul, "div": L.orc_ddiv}[op]
    rng = np.random.default_rng(7)
    a = edge_doubles(rng, 3000)
    b = edge_doubles(np.random.default_rng(8), 3000)
    b[:40] = a[:40]  # exact cancellations and quotients of 1
    checked = 0
    for x, y in zip(a, b):
        x, y = float(x), float(y)
        if op == "div" and y == 0:
            continue
        c = {"add": x + y, "sub": x - y, "mul": x * y, "div": x / y if y else 0.0}[op]
        if math.isinf(c) or math.isinf(PHI1 * abs(c) + ETA + abs(c)):
            continue  # overflow: both sides give inf - inf
        for up in (0, 1):
            assert bits(fn(E, up, x, y)) == bits(e_op(op, x, y, up)), (op, x, y, up)
        checked += 1
    assert checked > 3000


@pytest.mark.parametrize("name", SETS)
def test_recip_tables_are_the_cpu_division(name):
    o = po.Oracle(name)
    rd, ru = o.recip()
    for i, mi in enumerate(int(v) for v in o.moduli()):
        assert bits(float(rd[i])) == bits(e_op("div", 1.0, float(mi), False))
        assert bits(float(ru[i])) == bits(e_op("div", 1.0, float(mi), True))


# ---------------------------------------------------------------- evaluation, recomposed
def er_set_d(x):
    """src/extrange.cuh:46-58 for normal x: (significand in [1, 2), unbiased exponent); 0 -> (0, 0)"""
    if x == 0:
        return 0.0, 0
    fr, ex = math.frexp(x)
    return math.copysign(2 * abs(fr), x), ex - 1


class Eval:
    def __init__(self, o, policy):
        self.m = [int(v) for v in o.moduli()]
        self.N = len(self.m)
        self.M = exact.moduli_product(self.m)
        self.w = [pow(self.M // mi, -1, mi) for mi in self.m]
        c = o.eval_consts()
        self.acc, self.ref = c["acc"], c["ref_factor"]
        self.unit_low = (float(c["unit_low"]["frac"]), int(c["unit_low"]["exp"]))
        self.inv_unit_upp = (float(c["inv_unit_upp"]["frac"]), int(c["inv_unit_upp"]["exp"]))
        self.rrd = [e_op("div", 1.0, float(mi), False) for mi in self.m]
        self.rru = [e_op("div", 1.0, float(mi), True) for mi in self.m]
        self.pol = policy
        self.hits = set()

    def op(self, op, a, b, up):
        if self.pol == E:
            return e_op(op, a, b, up)
        q = {"add": Fraction(a) + Fraction(b), "sub": Fraction(a) - Fraction(b),
             "mul": Fraction(a) * Fraction(b)}[op]
        return exact.round_dir(q, up)

    def psum(self, v, up):
        def tree(t):
            if len(t) == 1:
                return t[0]
            h = len(t) // 2
            return self.op("add", tree(t[:h]), tree(t[h:]), up)
        return self.op("add", tree(v), 0.0, up)  # the root addition of 0 (psum(2N)<N>)

    def frac_part(self, s, up):
        w = float(int(s))
        return self.op("sub", s, w, up) if self.pol == D else s - w

    def fractions(self, s, up):
        if self.pol == D:
            return [self.op("mul", float(si), (self.rru if up else self.rrd)[i], up) for i, si in enumerate(s)]
        return [e_op("div", float(si), float(self.m[i]), up) for i, si in enumerate(s)]

    def __call__(self, digits, fast):
        d = [int(v) for v in digits]
        X = exact.crt(d, self.m)
        if self.pol == E and X == 0:
            self.hits.add("zero")
            return (0.0, 0), (0.0, 0)
        s = [(d[i] * self.w[i]) % mi for i, mi in enumerate(self.m)]
        # the first sums use the reciprocal tables in both policies
        fl = [self.op("mul", float(si), self.rrd[i], False) for i, si in enumerate(s)]
        fu = [self.op("mul", float(si), self.rru[i], True) for i, si in enumerate(s)]
        suml, sumu = self.psum(fl, False), self.psum(fu, True)
        if self.pol == D and suml == 0 and sumu == 0:
            self.hits.add("zero")
            return (0.0, 0), (0.0, 0)
        whl, whu = int(suml), int(sumu)
        suml, sumu = self.frac_part(suml, False), self.frac_part(sumu, True)
        low, upp = er_set_d(suml), er_set_d(sumu)
        if fast:
            if sumu >= self.acc:
                return low, upp
        else:
            mr = -1
            if whl != whu:
                mr = exact.mrc_digits(X, self.m)[-1]  # the last mixed-radix digit
            if mr > 0:
                self.hits.add("mr>0")
                return low, self.inv_unit_upp
            if mr == 0:
                self.hits.add("mr==0")
                low = self.unit_low
            if sumu >= self.acc:
                return low, upp
        K, steps = 0, 0
        while sumu < self.acc:
            t = -(math.ceil(math.log2(sumu)) + 1)
            assert t >= self.ref  # the REF_FACTOR clamp of MAX(t, REF_FACTOR) never fires
            k = max(t, self.ref)
            s = [(si * pow(2, k, mi)) % mi for si, mi in zip(s, self.m)]
            sumu = self.frac_part(self.psum(self.fractions(s, True), True), True)
            K += k
            steps += 1
        self.hits.add("refine1" if steps == 1 else "refine>1")
        suml = self.frac_part(self.psum(self.fractions(s, False), False), False)
        (lf, le), (uf, ue) = er_set_d(suml), er_set_d(sumu)
        return (lf, le - K), (uf, ue - K)


def residues(X, m):
    return [X % mi for mi in m]


def eval_inputs(ev, seed):
    """X over the whole range and at the branch-forcing ends: 0, tiny (multi-step refinement,
    mr == 0 ambiguity), X/M just under the accuracy constant (one step), near M (mr > 0)"""
    M = ev.M
    rng = random.Random(seed)
    acc = Fraction(ev.acc)
    xs = [0, 1, 2, 3, 5, 7, 2**40 + 1, 2**60 - 3, M - 1, M - 2, M - 3, M - 1000, M // 2, M // 2 + 1,
          M // 3, (M * 2) // 3]
    for f in (0.999, 0.9, 0.75, 0.51, 0.3, 1e-3, 1e-6, 2.0**-30, 2.0**-50, 2.0**-80):
        f = Fraction(f)
        xs += [int(f * acc * M), int(f * M), M - 1 - int(f * acc * M)]
    xs += [rng.randrange(M) for _ in range(300)]
    xs += [(1 << e) + rng.randrange(1 << min(e, 30)) for e in range(0, M.bit_length(), 7)]
    return [x % M for x in xs]


def ambiguous(ev, x):
    """whether the full evaluation of x takes an ambiguity branch (mr > 0 or mr == 0)"""
    saved = set(ev.hits)
    ev.hits.clear()
    ev(residues(x, ev.m), False)
    amb = bool({"mr>0", "mr==0"} & ev.hits)
    ev.hits = saved | ev.hits
    return amb


@pytest.mark.parametrize("name", SETS)
@pytest.mark.parametrize("policy", [D, E])
def test_eval_bits_recomposed(name, policy):
    o = po.Oracle(name)
    ev = Eval(o, policy)
    xs = eval_inputs(ev, 3)
    # the fast variant is defined for numbers without the ambiguity (its callers' numbers are
    # far below M); near M it would scale the wrapped value
    plain = [x for x in xs if not ambiguous(ev, x)]
    for fast, sel in ((False, xs), (True, plain)):
        digits = np.array([residues(x, ev.m) for x in sel], np.int32)
        lo, up = o.eval(digits, fast=fast, policy=policy)
        for i, x in enumerate(sel):
            wl, wu = ev(digits[i], fast)
            got = ((float(lo[i]["frac"]), int(lo[i]["exp"])), (float(up[i]["frac"]), int(up[i]["exp"])))
            assert (bits(got[0][0]), got[0][1], bits(got[1][0]), got[1][1]) == \
                (bits(wl[0]), wl[1], bits(wu[0]), wu[1]), (name, policy, fast, x)
    assert {"zero", "mr>0", "mr==0", "refine1", "refine>1"} <= ev.hits, ev.hits


@pytest.mark.parametrize("name", SETS)
@pytest.mark.parametrize("policy", [D, E])
def test_eval_fast_equals_full_and_width(name, policy):
    o = po.Oracle(name)
    ev = Eval(o, policy)
    xs = eval_inputs(ev, 4)
    n_all = len(xs)
    xs = [x for x in xs if not ambiguous(ev, x)]
    digits = np.array([residues(x, ev.m) for x in xs], np.int32)
    lf, uf = o.eval(digits, fast=False, policy=policy)
    lq, uq = o.eval(digits, fast=True, policy=policy)
    N = ev.N
    W = Fraction(N * (3 + int(math.log2(N)))) / 2**52 * (2 if policy == E else 1)
    same = 0
    for i, x in enumerate(xs):
        assert lf[i].tobytes() == lq[i].tobytes() and uf[i].tobytes() == uq[i].tobytes(), x
        same += 1
        if x == 0:
            continue
        lo, hi = exact.er_value(lf[i]), exact.er_value(uf[i])
        q = Fraction(x, ev.M)
        assert lo <= q <= hi
        if hi >= Fraction(ev.acc):  # no refinement: the absolute rounding-error bound
            assert hi - lo <= W, (x, float(hi - lo), float(W))
        else:  # refined: the relative accuracy the constant is built for (RNS_EVAL_RELATIVE_ERROR)
            assert hi - lo <= Fraction(2 if policy == E else 1, 10**7) * lo
    assert same > n_all // 2


@pytest.mark.parametrize("name", ["8_2double", "8_32bit"])
@pytest.mark.parametrize("policy", [D, E])
def test_set_d_and_round_use_full_and_fast_eval(name, policy):
    """mp_set_d evaluates with rns_eval_compute (src/arith/assign.cuh:48), mp_round and the round
    kernels with rns_eval_compute_fast (arith_utils.cuh:144, :190, mpvector.cuh:686): the eval
    bits of their results are those of the recomposed variant on the result's digits"""
    o = po.Oracle(name)
    ev = Eval(o, policy)
    x = o.set_d(gen.uexp(5, 200, -60, 60), policy)
    for r in x:
        (wl, wle), (wu, wue) = ev(r["digits"], False)
        assert (float(r["eval"][0]["frac"]), int(r["eval"][0]["exp"])) == (wl, wle)
        assert (float(r["eval"][1]["frac"]), int(r["eval"][1]["exp"])) == (wu, wue)
    # products of three 53-bit significands exceed these sets' precision (106, 120 bits): all round
    p = o.mul(o.mul(x, x[::-1].copy(), policy), x, policy)
    rounded = 0
    for r in p:
        (wl, wle), (wu, wue) = ev(r["digits"], True)
        assert (float(r["eval"][0]["frac"]), int(r["eval"][0]["exp"])) == (wl, wle)
        assert (float(r["eval"][1]["frac"]), int(r["eval"][1]["exp"])) == (wu, wue)
        rounded += 1
    assert rounded == len(p)
