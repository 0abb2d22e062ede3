"""mpres_dot_noalloc (the dot whose partial sums live in a fixed device array instead of an
allocation) against mpres_dot through the C ABI: bit-identical results on two calls in a
row, the second with a smaller input and reduction shape, and a clear error for a
reduction shape whose partials exceed the array."""
import ctypes as C
import importlib.util
import os

import numpy as np
import pytest

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


@pytest.mark.parametrize("name", ["8_2double", "32_8double"])
def test_dot_noalloc_matches_dot(name):
    c, o = mp.Context(name), po.Oracle(name)
    try:
        # (n, incx, incy, grid3, block3): the second call is smaller in every respect
        calls = [(200_000, 1, 1, 128, 64), (3001, -1, 2, 32, 32)]
        for i, (n, incx, incy, g3, b3) in enumerate(calls):
            print(f"[case] dot_noalloc {name} call {i}: n={n} inc=({incx},{incy}) shape={g3}x{b3}", flush=True)
            need = c.L.mpres_dot_workspace(c.h, n, g3, b3)
            assert need <= (4 << 20)
            lx, ly = 1 + (n - 1) * abs(incx), 1 + (n - 1) * abs(incy)
            xd, yd = gen.uexp(61 + i, lx, -20, 20), gen.uexp(71 + i, ly, -20, 20)
            ax, ay = c.array_set_d(xd), c.array_set_d(yd)
            r_ref, r_na = c.array(1), c.array(1)
            mp.check(c.L.mpres_dot(c.h, n, C.byref(ax), incx, C.byref(ay), incy, C.byref(r_ref), g3, b3, None, 0,
                                   None))
            mp.check(c.L.mpres_dot_noalloc(c.h, n, C.byref(ax), incx, C.byref(ay), incy, C.byref(r_na), g3, b3, None))
            got, want = c.array_to_host(r_na, 1), c.array_to_host(r_ref, 1)
            assert got.tobytes() == want.tobytes(), f"call {i}: noalloc {got} != dot {want}"
            # and both equal the oracle's dot in the same order
            orc = o.dot(o.set_d(xd, D), o.set_d(yd, D), g3, b3, D, incx=incx, incy=incy, n=n)
            assert got.tobytes() == orc.tobytes(), f"call {i}: noalloc {got} != oracle {orc}"
    finally:
        c.close()


def test_dot_noalloc_rejects_oversized_shape():
    c = mp.Context("32_8double")
    try:
        g3, b3 = 4096, 1024
        assert c.L.mpres_dot_workspace(c.h, 1 << 20, g3, b3) > (4 << 20)
        a = c.array_set_d(np.ones(16))
        r = c.array(1)
        rc = c.L.mpres_dot_noalloc(c.h, 16, C.byref(a), 1, C.byref(a), 1, C.byref(r), g3, b3, None)
        assert rc == -3  # MPRES_ERR_WORKSPACE
        assert b"at most" in mp.lib().mpres_last_error()
    finally:
        c.close()
