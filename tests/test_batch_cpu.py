"""The batched form on the CPU: the library's item layout (csrc/batch_plan.hpp)
and its persistent kernel (csrc/kernels_batch.hpp), compiled with g++ against
the host stand-in of the HIP header, one lane per number. The threads of a
workgroup run concurrently with a real barrier (tests/host/batch_emu.cpp).
Every problem's result must equal the oracle's single-call result (policy D)
with the same order arguments, at 8, 16 and 32 moduli, for any item size and
workgroup count."""
import ctypes as C
import importlib.util
import os
import subprocess

import numpy as np
import pytest

import gen
import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
D = po.POLICY_D

_spec = importlib.util.spec_from_file_location("mpres", os.path.join(ROOT, "mpres-blas_amd", "mpres.py"))
mp = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(mp)  # ctypes structs only; the HIP library is not loaded here


@pytest.fixture(scope="module")
def bemu(tmp_path_factory):
    out = tmp_path_factory.mktemp("bemu") / "libbemu.so"
    cmd = ["g++", "-std=c++17", "-O1", "-fPIC", "-shared", "-ffp-contract=off", "-pthread",
           "-I", os.path.join(ROOT, "tests", "host", "emu"),
           os.path.join(ROOT, "tests", "host", "batch_emu.cpp"),
           os.path.join(ROOT, "mpres-blas_amd", "csrc", "consts.cpp"), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    L = C.CDLL(str(out))
    vp = C.c_void_p
    L.emu_batch_workspace.restype = C.c_longlong
    L.emu_batch_workspace.argtypes = [C.c_int, vp, C.c_int, C.c_int]
    L.emu_batch_items.argtypes = [C.c_int, vp, C.c_int, C.c_int, vp, vp, vp, vp, C.c_int]
    L.emu_batch_run.argtypes = [C.c_int, vp, C.c_int, C.c_int, C.c_int, vp, C.c_size_t]
    return L


class HostArray:
    """mp_array_t over host memory (structure of arrays), from oracle records"""

    def __init__(self, o, recs):
        cnt = len(recs)
        self.o = o
        self.digits = np.ascontiguousarray(recs["digits"], np.int32)
        self.sign = np.ascontiguousarray(recs["sign"], np.int32)
        self.exp = np.ascontiguousarray(recs["exp"], np.int32)
        self.eval = np.zeros(2 * cnt, mp.ER_DTYPE)
        self.eval[:cnt] = recs["eval"][:, 0]
        self.eval[cnt:] = recs["eval"][:, 1]
        self.len = np.array([cnt], np.int32)
        self.s = mp.MpArray(digits=self.digits.ctypes.data, sign=self.sign.ctypes.data, exp=self.exp.ctypes.data,
                            eval=self.eval.ctypes.data, buf=None, len=self.len.ctypes.data)

    def records(self):
        cnt = int(self.len[0])
        out = self.o.empty(cnt)
        out["digits"] = self.digits
        out["sign"] = self.sign
        out["exp"] = self.exp
        out["eval"][:, 0] = self.eval[:cnt]
        out["eval"][:, 1] = self.eval[cnt:]
        return out


def _problems(o, n_mod):
    """a mixed batch: dots and gemvs of several shapes, strides, orders; returns
    (problems, host arrays, expected results per problem)"""
    lo, hi = (-20, 20) if n_mod == 8 else (-60, 60)
    keep, probs, want = [], [], []
    seed = [300]

    def arr(cnt):
        seed[0] += 1
        h = HostArray(o, o.set_d(gen.uexp(seed[0], cnt, lo, hi), D))
        keep.append(h)
        return h

    for n, g3, b3, incx, incy in [(1, 1, 1, 1, 1), (7, 2, 3, 1, 1), (300, 4, 64, 1, 1), (257, 3, 5, -1, 2),
                                  (1000, 1, 32, 2, -1), (64, 8, 8, 1, 1)]:
        x = arr(1 + (n - 1) * abs(incx))
        y = arr(1 + (n - 1) * abs(incy))
        r = arr(1)
        probs.append(mp.BatchProblem(op=mp.BATCH_DOT, n=n, incx=incx, incy=incy, grid3=g3, block3=b3,
                                     x=C.addressof(x.s), y=C.addressof(y.s), r=C.addressof(r.s)))
        want.append(("dot", r, o.dot(x.records(), y.records(), g3, b3, D, incx=incx, incy=incy, n=n)))
    for m, n, trans, b3, incx, incy in [(5, 7, 0, 1, 1, 1), (33, 20, 0, 3, 1, 1), (33, 20, 1, 8, 1, 1),
                                        (70, 9, 0, 2, -1, 2), (12, 40, 1, 1, 2, -1), (1, 1, 0, 4, 1, 1)]:
        lda = m + 2
        lx, ly = (m, n) if trans else (n, m)
        A = arr(lda * n)
        x = arr(1 + (lx - 1) * abs(incx))
        y = arr(1 + (ly - 1) * abs(incy))
        al, be = arr(1), arr(1)
        exp_y = o.gemv(trans, m, n, al.records(), A.records(), lda, x.records(), be.records(), y.records(), b3, D,
                       incx=incx, incy=incy)
        probs.append(mp.BatchProblem(op=mp.BATCH_GEMV, trans=mp.TRANS if trans else mp.NO_TRANS, m=m, n=n, lda=lda,
                                     incx=incx, incy=incy, block3=b3, alpha=C.addressof(al.s), A=C.addressof(A.s),
                                     x=C.addressof(x.s), beta=C.addressof(be.s), y=C.addressof(y.s)))
        want.append(("gemv", y, exp_y))
    # problems the single calls skip: nothing written
    z = arr(3)
    zr = arr(1)
    probs.append(mp.BatchProblem(op=mp.BATCH_DOT, n=0, incx=1, incy=1, grid3=2, block3=2, x=C.addressof(z.s),
                                 y=C.addressof(z.s), r=C.addressof(zr.s)))
    want.append(("dot", zr, zr.records().copy()))
    return probs, keep, want


@pytest.mark.parametrize("name", ["8_2double", "16_4double", "32_8double"])
@pytest.mark.parametrize("rows,wgs", [(1000, 1), (16, 3), (5, 2)])
def test_batch_kernel_matches_single_calls(bemu, name, rows, wgs):
    o = po.Oracle(name)
    sid = po.set_names().index(name)
    probs, keep, want = _problems(o, len(o.moduli()))
    arr = (mp.BatchProblem * len(probs))(*probs)
    wsb = bemu.emu_batch_workspace(sid, arr, len(probs), rows)
    assert wsb > 0
    ws = np.zeros(wsb, np.uint8)
    assert bemu.emu_batch_run(sid, arr, len(probs), rows, wgs, ws.ctypes.data, wsb) == 0
    for i, (kind, h, exp) in enumerate(want):
        got = h.records()
        assert got.tobytes() == np.asarray(exp).tobytes(), (name, rows, wgs, i, kind)


def test_batch_layout(bemu):
    """items: every gemv split into runs of rows covering its outputs once, launch order by
    descending work, scratch ranges aligned and disjoint; skipped problems give no item"""
    o = po.Oracle("16_4double")
    sid = po.set_names().index("16_4double")
    probs, keep, want = _problems(o, 16)
    arr = (mp.BatchProblem * len(probs))(*probs)
    rows = 8
    cap = 256
    kind, r0, r1 = (np.zeros(cap, np.int32) for _ in range(3))
    off = np.zeros(cap, np.int64)
    cnt = bemu.emu_batch_items(sid, arr, len(probs), rows, kind.ctypes.data, r0.ctypes.data, r1.ctypes.data,
                               off.ctypes.data, cap)
    ndot = sum(1 for p in probs if p.op == mp.BATCH_DOT and p.n > 0)
    ngemv = sum(-(-(p.n if p.trans != mp.NO_TRANS else p.m) // rows) for p in probs if p.op == mp.BATCH_GEMV)
    assert cnt == ndot + ngemv
    assert (kind[:cnt] == 0).sum() == ndot
    assert all(0 <= a < b and b - a <= rows for a, b, k in zip(r0[:cnt], r1[:cnt], kind[:cnt]) if k != 0)
    assert all(v % 256 == 0 for v in off[:cnt])
    assert len(set(off[:cnt].tolist())) == cnt
    wsb = bemu.emu_batch_workspace(sid, arr, len(probs), rows)
    assert wsb >= off[:cnt].max()


def test_batch_rejects_bad_problems(bemu):
    o = po.Oracle("8_2double")
    sid = po.set_names().index("8_2double")
    x = HostArray(o, o.set_d(gen.uexp(1, 4, -3, 3), D))
    bad = [mp.BatchProblem(op=mp.BATCH_DOT, n=4, incx=1, incy=1, grid3=0, block3=4, x=C.addressof(x.s),
                           y=C.addressof(x.s), r=C.addressof(x.s)),
           mp.BatchProblem(op=mp.BATCH_GEMV, trans=mp.NO_TRANS, m=2, n=2, lda=1, incx=1, incy=1, block3=1,
                           alpha=C.addressof(x.s), A=C.addressof(x.s), x=C.addressof(x.s), beta=C.addressof(x.s),
                           y=C.addressof(x.s)),
           mp.BatchProblem(op=7, n=4)]
    for p in bad:
        arr = (mp.BatchProblem * 1)(p)
        assert bemu.emu_batch_workspace(sid, arr, 1, 16) == -1
    assert bemu.emu_batch_workspace(sid, (mp.BatchProblem * 1)(bad[0]), 1, 0) == -1
