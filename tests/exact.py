"""Exact (big-integer / rational) arithmetic used to pin the oracle and to
measure errors. Independent of both the oracle and the HIP library: every
quantity here is recomputed from its mathematical definition.
"""
import math
from fractions import Fraction


def moduli_product(m):
    M = 1
    for v in m:
        M *= int(v)
    return M


def crt(digits, m):
    """Integer X in [0, M) with X = digits[i] mod m[i]."""
    M = moduli_product(m)
    X = 0
    for d, mi in zip(digits, m):
        mi = int(mi)
        Mi = M // mi
        X += int(d) * Mi * pow(Mi, -1, mi)
    return X % M


def mp_value(rec, m):
    """Exact value of one mp_float_t record (numpy void with digits/sign/exp)."""
    X = crt(rec["digits"], m)
    v = Fraction(X) * (Fraction(2) ** int(rec["exp"]))
    return -v if int(rec["sign"]) else v


def er_value(e):
    return Fraction(float(e["frac"])) * (Fraction(2) ** int(e["exp"]))


def dbl_exact(x):
    return Fraction(float(x))


def round_dir(q, up):
    """Directed rounding of a rational to binary64 (normal range)."""
    f = float(q)  # correctly rounded to nearest
    fq = Fraction(f)
    if fq == q:
        return f
    if up:
        return f if fq > q else math.nextafter(f, math.inf)
    return f if fq < q else math.nextafter(f, -math.inf)


def round_dir_er(q, up):
    """Directed rounding of a positive rational to (frac in [0.5,1), exp), as mpfr_get_d_2exp."""
    e = q.numerator.bit_length() - q.denominator.bit_length()
    scaled = q / (Fraction(2) ** e)  # in (0.5, 2)
    f = round_dir(scaled, up)
    fr, ex = math.frexp(f)
    return fr, ex + e


def mrc_digits(X, m):
    """Mixed-radix digits of X: X = d0 + d1*m0 + d2*m0*m1 + ..."""
    out = []
    for mi in m:
        out.append(X % int(mi))
        X //= int(mi)
    return out


def exact_rank(digits, m, w):
    """k with sum_i M_i * (x_i w_i mod m_i) = X + k*M, i.e. floor(sum s_i/m_i)."""
    s = [(int(d) * int(wi)) % int(mi) for d, wi, mi in zip(digits, w, m)]
    tot = sum(Fraction(si, int(mi)) for si, mi in zip(s, m))
    return math.floor(tot)


def residues(X, m):
    return [X % int(mi) for mi in m]
