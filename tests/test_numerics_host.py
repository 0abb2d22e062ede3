"""The GPU path's scalar numerics (mpres-blas_amd/csrc/numerics.hpp), built
for the host, against the oracle's IEEE directed rounding (fesetround), glibc's
ceil(log2()) and exact integer arithmetic."""
import ctypes as C
import math
import os
import struct
from fractions import Fraction
import subprocess
import tempfile

import numpy as np
import pytest

import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def bits(x):
    return struct.unpack("<Q", struct.pack("<d", x))[0]


@pytest.fixture(scope="module")
def nh():
    d = tempfile.mkdtemp()
    so = os.path.join(d, "nh.so")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-o", so,
                    os.path.join(ROOT, "tests", "host", "numerics_host.cpp"),
                    os.path.join(ROOT, "mpres-blas_amd", "csrc", "consts.cpp")], check=True)
    L = C.CDLL(so)
    for f in ("nh_add", "nh_sub", "nh_mul", "nh_div", "nh_divk"):
        getattr(L, f).restype = C.c_double
        getattr(L, f).argtypes = [C.c_int, C.c_double, C.c_double]
    L.nh_unit.argtypes = [C.c_int, C.POINTER(C.c_double)]
    L.nh_er_adjust.argtypes = [C.c_double, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.nh_ceil_log2.argtypes = [C.c_double]
    L.nh_mulmod.argtypes = [C.c_int, C.c_int, C.c_int]
    L.nh_modd.argtypes = [C.c_double, C.c_int]
    u32p, ip = C.POINTER(C.c_uint32), C.POINTER(C.c_int)
    L.nh_big_shl.argtypes = L.nh_big_shr.argtypes = [u32p, C.c_int, u32p]
    L.nh_big_bits.argtypes = [u32p, C.c_int, ip]
    L.nh_big_cmp.argtypes = [u32p, u32p]
    L.nh_big_round.argtypes = [u32p, C.c_int, u32p]
    L.nh_big_mod.argtypes = [u32p, C.c_int, ip]
    L.nh_er_add.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_double),
                            C.POINTER(C.c_int)]
    return L


def operands(rng, count):
    u = rng.uniform(-1, 1, count)
    e = rng.integers(-80, 80, count)
    a = np.ldexp(u, e)
    out = list(a)
    # edge cases: zeros of both signs, exact cancellation, powers of two, ulp neighbours
    out += [0.0, -0.0, 1.0, -1.0, 2.0**-60, 3.0, 1.5, math.nextafter(1.0, 2), math.nextafter(1.0, 0)]
    return out


@pytest.mark.parametrize("op", ["add", "sub", "mul", "div"])
def test_directed_ops_match_ieee(nh, oracle_lib, op):
    L = oracle_lib.lib()
    ref = {"add": L.orc_dadd, "sub": L.orc_dsub, "mul": L.orc_dmul, "div": L.orc_ddiv}[op]
    mine = {"add": nh.nh_add, "sub": nh.nh_sub, "mul": nh.nh_mul, "div": nh.nh_div}[op]
    rng = np.random.default_rng(11)
    xs = operands(rng, 4000)
    ys = operands(rng, 4000)
    ys[:200] = [-v for v in xs[:200]]  # exact cancellation
    ys[200:400] = xs[200:400]
    for x, y in zip(xs, ys):
        if op == "div" and y == 0:
            continue
        for up in (0, 1):
            a, b = mine(up, x, y), ref(po.POLICY_D, up, x, y)
            assert bits(a) == bits(b), (op, up, x, y, a, b)


def test_ceil_log2_matches_libm(nh, oracle_lib):
    L = oracle_lib.lib()
    rng = np.random.default_rng(12)
    xs = []
    for e in list(range(-1020, -1)) + list(range(1, 60)):
        base = 2.0**e
        xs.append(base)
        v = base
        for _ in range(12):  # the first doubles above a power of two
            v = math.nextafter(v, math.inf)
            xs.append(v)
        # around the rounding threshold of log2 near 2^e
        for j in range(1, 600, 7):
            xs.append(base * (1 + j * 2.0**-52))
        xs.append(math.nextafter(base, 0))
    xs += list(np.ldexp(rng.uniform(0.5, 1, 3000), rng.integers(-1000, -20, 3000)))
    bad = [x for x in xs if nh.nh_ceil_log2(x) != L.orc_ceil_log2_lib(x)]
    assert not bad, bad[:5]


@pytest.mark.parametrize("m", [113812103, 99312287, 1283742851, 1122512269, 39051, 34571])
def test_mulmod_exact(nh, m):
    rng = np.random.default_rng(m)
    vals = list(rng.integers(0, m, 3000)) + [0, 1, m - 1, m - 2, m // 2, (m + 1) // 2]
    ws = list(rng.integers(0, m, 3000)) + [m - 1, m - 1, m - 1, 2, 2, 2]
    for a, b in zip(vals, ws):
        assert nh.nh_mulmod(int(a), int(b), m) == (int(a) * int(b)) % m
    # products that land near multiples of m (where a quotient estimate is off by one)
    for k in range(1, 3000):
        a = int(rng.integers(1, m))
        if math.gcd(a, m) != 1:
            continue
        b = (k * m + int(rng.integers(-3, 4))) * pow(a, -1, m) % m
        assert nh.nh_mulmod(a, b, m) == (a * b) % m
    for v in list(rng.integers(-2**52, 2**52, 3000)) + [0, 2**53 - 1, -(2**53 - 1)]:
        assert nh.nh_modd(float(v), m) == int(v) % m


def test_division_by_unit_matches_ieee(nh, oracle_lib):
    """divk_rd / divk_ru (reciprocal of the fixed divisor in two doubles, exact remainder)
    against IEEE directed division: the actual unit fractions of every moduli set (and the
    reciprocals the constants store), random divisors, dividends that are products of two
    fractions (as in er_md), and quotients that are exact or within one rounding of a double"""
    L = oracle_lib.lib()
    rng = np.random.default_rng(23)
    zs = []
    for sid in range(len(po.set_names())):
        u = (C.c_double * 6)()
        assert nh.nh_unit(sid, u) == 0
        zs += [u[0], u[1]]
        for z, zh, zl in ((u[0], u[2], u[3]), (u[1], u[4], u[5])):
            e = 1 - Fraction(zh) * Fraction(z)  # what the fma forms: exact and representable
            assert zh == 1.0 / z and Fraction(float(e)) == e and zl == float(e) / z
    zs += list(rng.uniform(0.5, 1.0, 40)) + list(rng.uniform(1.0, 2.0, 40)) + [0.5, 1.0, math.nextafter(1.0, 0)]
    checked = 0
    for z in zs:
        fx = rng.uniform(1.0, 2.0, 300)
        fy = rng.uniform(1.0, 2.0, 300)
        a = [x * y for x, y in zip(fx, fy)]
        a += list(rng.uniform(1.0, 4.0, 100)) + [0.0, -0.0, 1.0, 2.0, math.nextafter(4.0, 0)]
        q = rng.uniform(0.5, 8.0, 200)
        a += [float(v * z) for v in q]  # a/z within a rounding of a double (often exact)
        for v in a:
            for up in (0, 1):
                got, want = nh.nh_divk(up, v, z), L.orc_ddiv(po.POLICY_D, up, v, z)
                assert bits(got) == bits(want), (up, v, z, got, want)
                checked += 1
    assert checked > 100000


def test_er_adjust_matches_bit_definition(nh):
    """er_adjust through frexp against the reference's definition on the bit pattern
    (exponent field read out, mantissa given the exponent of 1.0): normal inputs of both
    signs over a wide range, powers of two and their neighbours, and signed zeros (the
    inputs er_adjust receives)"""
    def ref(f, e):
        u = bits(f)
        ex = (e + ((u & ~(1 << 63)) >> 52) - 1023) if f != 0 else 0
        fb = (u & (((1 << 52) - 1) | (1 << 63))) | (1023 << 52)
        return struct.unpack("<d", struct.pack("<Q", fb))[0] * (1.0 if f != 0 else 0.0), ex
    rng = np.random.default_rng(5)
    fs = list(np.ldexp(rng.uniform(-2, 2, 20000), rng.integers(-300, 300, 20000)))
    fs += [0.0, -0.0, 1.0, -1.0, 2.0, 0.5, 3.999999999999999, math.nextafter(1.0, 0), math.nextafter(2.0, 0),
           math.nextafter(4.0, 0), 2.0**-1000, -2.0**1000]
    rf, re_ = C.c_double(), C.c_int()
    for i, f in enumerate(fs):
        e = int(rng.integers(-2000, 2000))
        nh.nh_er_adjust(float(f), e, C.byref(rf), C.byref(re_))
        wf, we = ref(float(f), e)
        assert bits(rf.value) == bits(wf) and re_.value == we, (f, e, rf.value, re_.value, wf, we)


BW = 24  # limbs of the host build's fixed-width integers


def _big(v):
    a = np.array([(v >> (32 * i)) & 0xffffffff for i in range(BW)], np.uint32)
    return a, a.ctypes.data_as(C.POINTER(C.c_uint32))


def _int(a):
    return sum(int(x) << (32 * i) for i, x in enumerate(a))


def _big_values(rng, count):
    """integers of every length up to 24 limbs, sparse and dense bit patterns, single bits,
    all-ones runs, zero"""
    out = [0, 1, (1 << (32 * BW)) - 1, 1 << (32 * BW - 1), (1 << 32) - 1, 1 << 32]
    for _ in range(count):
        bl = int(rng.integers(1, 32 * BW + 1))
        kind = int(rng.integers(0, 3))
        if kind == 0:
            v = int.from_bytes(rng.bytes(BW * 4), "little") >> (32 * BW - bl)
        elif kind == 1:
            v = (1 << bl) - 1 - ((1 << int(rng.integers(0, bl))) - 1)
        else:
            v = 1 << (bl - 1) | (1 << int(rng.integers(0, bl)))
        out.append(v)
    return out


def test_bigint_shifts_and_bits(nh):
    """big_shl / big_shr (barrel of whole limbs, then bits), bit length, trailing zeros, single
    bits, the sticky test and comparison, against Python integers at shifts across and beyond
    the width"""
    rng = np.random.default_rng(21)
    mask = (1 << (32 * BW)) - 1
    out = np.zeros(BW, np.uint32)
    op = out.ctypes.data_as(C.POINTER(C.c_uint32))
    bits4 = np.zeros(4, np.int32)
    vals = _big_values(rng, 400)
    for t, v in enumerate(vals):
        a, ap = _big(v)
        for s in [0, 1, 31, 32, 33, 63, 64, 100, 255, 256, 500, 767, 768, 800, 1000, int(rng.integers(0, 800))]:
            nh.nh_big_shl(ap, s, op)
            assert _int(out) == (v << s) & mask, (v, s)
            nh.nh_big_shr(ap, s, op)
            assert _int(out) == v >> s, (v, s)
        i = int(rng.integers(-2, 32 * BW + 3))
        nh.nh_big_bits(ap, i, bits4.ctypes.data_as(C.POINTER(C.c_int)))
        tz = (v & -v).bit_length() - 1 if v else 0
        bit = (v >> i) & 1 if 0 <= i < 32 * BW else 0
        below = int(i > 0 and (v & ((1 << i) - 1)) != 0)
        assert list(bits4) == [v.bit_length(), tz, bit, below], (v, i, list(bits4))
        w = vals[(t * 7 + 3) % len(vals)] if t % 3 else v
        b, bp = _big(w)
        assert nh.nh_big_cmp(ap, bp) == (v > w) - (v < w)


def test_bigint_round_and_mod(nh):
    """big_round (to nearest, ties to even, the exponent growing by the bits dropped) and
    a mod m through FP64 steps (big_mod_d) and through integer division (big_mod_u32),
    against Python integers; ties are built on purpose"""
    rng = np.random.default_rng(22)
    out = np.zeros(BW, np.uint32)
    op = out.ctypes.data_as(C.POINTER(C.c_uint32))
    r2 = np.zeros(2, np.int32)
    vals = _big_values(rng, 300)
    for p in (53, 106, 212):
        vals += [((1 << p) + 1) << 5 | 1 << 4, ((1 << p) | 1) << 7 | 1 << 6, ((1 << p) - 1) << 3 | 1 << 2]
    moduli = [2147483647, 2147483629, 1073741789, 65521, 3, 2, 1000003, 2147418113]
    for v in vals:
        a, ap = _big(v)
        for p in (53, 106, 212):
            e = nh.nh_big_round(ap, p, op)
            sh = max(v.bit_length() - p, 0)
            q, low = v >> sh, v & ((1 << sh) - 1)
            if sh and (low > 1 << (sh - 1) or (low == 1 << (sh - 1) and q & 1)):
                q += 1
            assert (_int(out), e) == (q, sh), (v, p)
        for m in moduli:
            nh.nh_big_mod(ap, m, r2.ctypes.data_as(C.POINTER(C.c_int)))
            assert list(r2) == [v % m, v % m], (v, m)

