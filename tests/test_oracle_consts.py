"""Pins the oracle's RNS / mp constants (restating src/rns.cuh:324-442 and
src/arith/arith_utils.cuh:44-85) against an independent exact recomputation."""
import math
from fractions import Fraction

import numpy as np
import pytest

import exact
import pyoracle as po

SETS = ["8_2double", "16_4double", "32_8double", "8_32bit", "16_32bit", "32_32bit", "8_16bit", "16_16bit", "32_16bit"]

# MP_PRECISION / MP_H / MP_J and RNS_EVAL_ACCURACY / REF_FACTOR of the moduli sets
# the measured configs use, as derived independently with exact integers.
KNOWN = {
    "8_2double": (106, -108, -1, 2.13e-7, 21),
    "16_4double": (212, -214, -1, 5.68e-7, 19),
    "32_8double": (424, -426, -1, 1.42e-6, 18),
    "8_32bit": (120, -122, -1, 2.13e-7, 21),
}


@pytest.fixture(scope="module", params=SETS)
def orc(request, oracle_lib):
    return po.Oracle(request.param)


def test_moduli_and_log2(orc):
    m = [int(v) for v in orc.moduli()]
    M = exact.moduli_product(m)
    # RNS_MODULI_PRODUCT_LOG2 in the params files is floor(log2 M)
    assert orc.log2m == M.bit_length() - 1
    for i in range(len(m)):
        for j in range(i + 1, len(m)):
            assert math.gcd(m[i], m[j]) == 1


def test_part_inverse(orc):
    m = [int(v) for v in orc.moduli()]
    M = exact.moduli_product(m)
    w = orc.part_inverse()
    for i, mi in enumerate(m):
        assert ((M // mi) * int(w[i])) % mi == 1
        assert 0 <= w[i] < mi


def test_pow2_tables(orc):
    m = [int(v) for v in orc.moduli()]
    M = exact.moduli_product(m)
    P = orc.pow2_table()
    for k in (0, 1, 2, 29, 30, 31, 100 % (orc.log2m + 1), orc.log2m):
        assert [pow(2, k, mi) for mi in m] == list(P[k])
    H = orc.part_pow2_residues()
    Mp = orc.pow2_residues()
    Pinv = orc.pow2_inverse()
    for j in range(30):
        mod = 2 ** (j + 1)
        assert Mp[j] == M % mod
        assert list(H[j]) == [(M // mi) % mod for mi in m]
        assert [(int(Pinv[j][i]) * pow(2, j + 1, mi)) % mi for i, mi in enumerate(m)] == [1] * len(m)


def test_mrc_inverse(orc):
    m = [int(v) for v in orc.moduli()]
    T = orc.mrc_inverse()
    for i in range(len(m)):
        for j in range(len(m)):
            if j > i:
                assert (int(T[i][j]) * m[i]) % m[j] == 1
            else:
                assert T[i][j] == 0


def test_eval_units(orc):
    m = [int(v) for v in orc.moduli()]
    M = exact.moduli_product(m)
    # M rounded down to 53 significant bits (MPFR default precision, MPFR_RNDD)
    s = M.bit_length() - 53
    M53 = (M >> s) << s
    c = orc.eval_consts()
    inv = Fraction(1, M53)
    for key, up in (("unit_low", False), ("unit_upp", True)):
        fr, ex = exact.round_dir_er(inv, up)
        assert (float(c[key]["frac"]), int(c[key]["exp"])) == (fr, ex), key
        assert 0.5 <= float(c[key]["frac"]) < 1.0
    low_u = exact.er_value(c["unit_low"])
    upp_u = exact.er_value(c["unit_upp"])
    assert low_u <= inv <= upp_u
    fr, ex = exact.round_dir_er(1 - upp_u, True)
    assert (float(c["inv_unit_upp"]["frac"]), int(c["inv_unit_upp"]["exp"])) == (fr, ex)
    fr, ex = exact.round_dir_er(1 - low_u, False)
    assert (float(c["inv_unit_low"]["frac"]), int(c["inv_unit_low"]["exp"])) == (fr, ex)


def test_accuracy_and_mp_consts(orc):
    m = [int(v) for v in orc.moduli()]
    M = exact.moduli_product(m)
    n = len(m)
    c = orc.eval_consts()
    acc = c["acc"]
    # 4 u n log2(n) (1 + eps/2) / eps with u = 2^-53, eps = 1e-7: checked to 1e-15 relative
    want = Fraction(4 * n * int(math.log2(n))) / Fraction(2**52) * (1 + Fraction(1, 2 * 10**7)) * 10**7
    assert abs(Fraction(acc) - want) / want < Fraction(1, 10**15)
    assert c["ref_factor"] == math.floor(math.log2(1 / (2 * acc)))
    prec, h, j = orc.mp_consts()
    s = M.bit_length() - 53
    M53 = (M >> s) << s
    # 2^MP_H <= sqrt(M53)/M53 < 2^(MP_H+1) ; MP_PRECISION = floor(log2 sqrt(M53)) - 1
    assert Fraction(2) ** h <= 1 / Fraction(math.isqrt(M53)) < Fraction(2) ** (h + 2)
    assert math.isqrt(M53).bit_length() - 1 - 1 == prec
    assert 2 ** (-2 * h - 2) <= M53 < 2 ** (-2 * h)
    assert j == -1
    name = orc.set_name
    if name in KNOWN:
        kp, kh, kj, kacc, kref = KNOWN[name]
        assert (prec, h, j, c["ref_factor"]) == (kp, kh, kj, kref)
        assert abs(acc - kacc) / kacc < 5e-3


def test_recip(orc):
    m = [int(v) for v in orc.moduli()]
    rd, ru = orc.recip()
    for i, mi in enumerate(m):
        q = Fraction(1, mi)
        assert Fraction(float(rd[i])) < q < Fraction(float(ru[i]))
        # emulated rounding widens by at most two ulps
        rn = float(q)
        assert math.nextafter(math.nextafter(rn, -1), -1) <= rd[i] and ru[i] <= math.nextafter(math.nextafter(rn, 2), 2)
