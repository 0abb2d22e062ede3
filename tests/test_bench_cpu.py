"""bench.py's CPU-side helpers: the vectorised exact dot product used for the error figure
of the 2^24-term dot, against the term-by-term exact sum."""
import os
import sys

import numpy as np
import pytest

import gen

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


@pytest.mark.parametrize("n,lo,hi,seed", [(1, -5, 5, 1), (4096, -20, 20, 2), (30000, -300, 300, 3),
                                          (20000, -1000, 1000, 4)])
def test_exact_dot_fast_matches_exact_dot(n, lo, hi, seed):
    a, b = gen.uexp(seed, n, lo, hi), gen.uexp(seed + 100, n, lo, hi)
    a[::7] = 0.0
    b[::11] = -b[::11]
    assert bench.exact_dot_fast(a, b) == bench.exact_dot(a, b)


def test_exact_dot_fast_cancellation_and_zero():
    a = np.array([1.5, 1.5, 2.0 ** -60, 0.0])
    b = np.array([0.75, -0.75, 3.0, 5.0])
    assert bench.exact_dot_fast(a, b) == bench.exact_dot(a, b) == 3 * bench.Fraction(2) ** -60
    assert bench.exact_dot_fast(np.zeros(3), np.ones(3)) == 0
