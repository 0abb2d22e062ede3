"""The library's reduction kernels, compiled for the CPU with g++ (and with clang,
locals initialised to a non-zero pattern) against a host stand-in of the HIP header
(tests/host/emu) and run one lane per number, against the oracle. This checks the kernels' own logic - the fast kernels'
abandoning and marking of rare chains, the fixup kernels, the gemm epilogues -
on the CPU, at 8, 16 and 32 moduli, including chains that need rounding and
exact cancellations."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import gen
import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
D = po.POLICY_D
SETS = {"8_2double": (-20, 20), "16_4double": (-150, 150), "32_8double": (-400, 400)}


CLANG = "/opt/rocm/llvm/bin/clang++"
BUILDS = {
    "g++": ["g++", "-O1"],
    # every local variable starts as a non-zero byte pattern: a read of an uninitialised
    # local (which the GPU would see as a stale register) changes the results
    "clang-pattern": [CLANG, "-O2", "-ftrivial-auto-var-init=pattern", "-Wno-unknown-warning-option"],
}


@pytest.fixture(scope="module", params=[b for b in BUILDS if b == "g++" or os.path.exists(CLANG)])
def emu(request, tmp_path_factory):
    out = tmp_path_factory.mktemp("emu") / "libkemu.so"
    # one lane per thread: the SpMV residue kernel's term slots (combined by shuffles) become one
    cmd = BUILDS[request.param] + ["-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-pthread", "-DMPH_SP_TS=1", "-DMPH_SPW_T=1",
           "-I", os.path.join(ROOT, "tests", "host", "emu"),
           os.path.join(ROOT, "tests", "host", "kernels_emu.cpp"),
           os.path.join(ROOT, "mpres-blas_amd", "csrc", "consts.cpp"), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return C.CDLL(str(out))


def P(a):
    return a.ctypes.data_as(C.c_void_p)


def _gemm_inputs(o, lo, hi, m, n, k, cancel=True):
    As = gen.uexp(91, m * k, lo, hi).reshape(k, m)
    Bs = gen.uexp(92, k * n, lo, hi).reshape(n, k)
    if cancel:  # exact cancellations between consecutive terms (every entry)
        for i in range(0, min(k - 1, 8), 2):
            As[i + 1] = As[i]
            Bs[:, i + 1] = -Bs[:, i]
    else:  # one column of B with a cancelling pair: only those entries take the fixup path
        As[1] = As[0]
        Bs[2, 1] = -Bs[2, 0]
    A, B = o.set_d(As.ravel(), D), o.set_d(Bs.ravel(), D)
    Cm = o.set_d(gen.uexp(93, m * n, lo, hi), D)
    return A, B, Cm


@pytest.mark.parametrize("name", list(SETS))
@pytest.mark.parametrize("k", [1, 2, 5, 24])
def test_gemm_v2_fast_fixup_epilogue(emu, name, k):
    o = po.Oracle(name)
    sid = po.set_names().index(name)
    lo, hi = SETS[name]
    m, n = 9, 8
    A, B, Cm = _gemm_inputs(o, lo, hi, m, n, k)
    alpha, beta = o.set_d([0.5], D), o.set_d([3.0], D)
    want = o.gemm_v2(0, 0, m, n, k, alpha, A, m, B, k, beta, Cm, m, D)
    for split in (1, 0):
        got, ws = Cm.copy(), o.empty(m * n)
        assert emu.emu_gemm_v2(sid, 0, 0, m, n, k, P(alpha), P(A), m, P(B), k, P(beta), P(got), m, P(ws), split) == 0
        assert got.tobytes() == want.tobytes(), (name, k, split)
        assert not (ws["sign"] == 2).any()  # every marked sum was recomputed


@pytest.mark.parametrize("name", list(SETS))
def test_dot_pass1_fast_fixup(emu, name):
    """First dot pass (strided partial sums) with rounding and cancellations,
    against the oracle's dot with one block (g3 = 1), whose first-pass partials
    are then reduced by the same tree."""
    o = po.Oracle(name)
    sid = po.set_names().index(name)
    lo, hi = SETS[name]
    cnt, nvt = 600, 32
    xs, ys = gen.uexp(71, cnt, lo, hi), gen.uexp(72, cnt, lo, hi)
    idx = np.arange(0, cnt - nvt, 3 * nvt)
    xs[idx + nvt], ys[idx + nvt] = xs[idx], -ys[idx]
    x, y = o.set_d(xs, D), o.set_d(ys, D)
    for split in (1, 0):
        part = o.empty(nvt)
        assert emu.emu_dot_v2_pass1(sid, cnt, nvt, P(x), P(y), P(part), split) == 0
        # partial v = sum over i = v, v + nvt, ... of x_i * y_i, in order, from zero
        for v in range(nvt):
            acc = o.set_d([0.0], D)
            for i in range(v, cnt, nvt):
                acc = o.add(acc, o.mul(x[i:i + 1], y[i:i + 1], D), D)
            assert part[v].tobytes() == acc[0].tobytes(), (name, split, v)


@pytest.mark.parametrize("name,lo,hi", [("16_4double", -30, 30), ("16_4double", -150, 150), ("32_8double", -30, 30),
                                        ("32_8double", -400, 400), ("16_32bit", -40, 40), ("32_32bit", -40, 40)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("cancel", [False, True])
def test_gemm_v2_split_form(emu, name, lo, hi, ta, tb, cancel):
    """Split form: scalar chains + pre-scaled residue products + digit reconstruction,
    with zeros, cancellations (marked and recomputed) and every transpose."""
    o = po.Oracle(name)
    sid = po.set_names().index(name)
    m, n, k = 7, 6, 13
    A, B, Cm = _gemm_inputs(o, lo, hi, m, n, k, cancel)
    z = o.set_d([0.0, -0.0], D)
    A[3], B[5] = z[0], z[1]
    alpha, beta = o.set_d([0.5], D), o.set_d([-3.0], D)
    lda = k if ta else m
    ldb = n if tb else k
    want = o.gemm_v2(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, Cm, m, D)
    got, ws = Cm.copy(), o.empty(m * n)
    assert emu.emu_gemm_v2_split(sid, ta, tb, m, n, k, P(alpha), P(A), lda, P(B), ldb, P(beta), P(got), m, P(ws)) == 0
    assert got.tobytes() == want.tobytes()


@pytest.mark.parametrize("name,lo,hi", [("16_4double", -30, 30), ("16_4double", -200, 200), ("32_8double", -30, 30),
                                        ("32_8double", -400, 400), ("16_32bit", -40, 40)])
def test_spmv_split_form(emu, name, lo, hi):
    """Split-form CSR SpMV (scalar chains, residues against the row exponent, fixup of the
    marked rows) on a power-law matrix with zero values, zero x entries, cancelling
    pairs and wide exponents, against the oracle's CSR kernel"""
    o = po.Oracle(name)
    sid = po.set_names().index(name)
    m = 400
    irp, ja, as_ = gen.powerlaw_csr(81, m, kmax=300)
    assert (np.diff(irp) >= 128).sum() >= 5  # rows for the long-row residue kernel
    as_[::13] = 0.0
    xs = gen.uexp(82, m, lo, hi)
    xs[::17] = 0.0
    # rows whose first two terms cancel exactly
    for r in range(0, m, 9):
        a, b = irp[r], irp[r + 1]
        if b - a >= 2:
            ja[a + 1] = ja[a]
            as_[a + 1] = -as_[a]
    x = o.set_d(xs, D)
    want = o.spmv(m, irp, ja, as_, x, 1, D)
    got = o.empty(m)
    counts = np.zeros(2, np.int32)
    assert emu.emu_spmv_split(sid, m, P(np.ascontiguousarray(irp, np.int32)), P(np.ascontiguousarray(ja, np.int32)),
                              P(np.ascontiguousarray(as_)), P(x), P(got), P(counts)) == 0
    assert got.tobytes() == want.tobytes(), name
    # the cancelling pairs reach the fixup; with narrow exponents long rows reach their kernel
    assert counts[1] > 0 and (hi > 100 or counts[0] > 0), counts


@pytest.mark.parametrize("name", ["16_4double", "32_8double"])
@pytest.mark.parametrize("ncols_of", [lambda m: 3 * m, lambda m: m // 2], ids=["wide", "tall"])
def test_spmv_split_rectangular(emu, name, ncols_of):
    """Split-form CSR SpMV on a rectangular matrix: x has ncols records. The gathers read the
    packed copy of x for the columns below min(m, largest column + 1) and x itself for the
    others (wide); a tall matrix's x has fewer records than the matrix has rows, and only the
    columns it uses are packed"""
    o = po.Oracle(name)
    sid = po.set_names().index(name)
    m = 300
    ncols = ncols_of(m)
    irp, ja, as_ = gen.powerlaw_csr(83, m, kmax=600)
    rng = np.random.default_rng(84)
    ja = rng.integers(0, ncols, ja.size).astype(np.int32)  # CSR order kept, columns unsorted
    ja[irp[1]] = ncols - 1  # the largest column is used
    as_[::11] = 0.0
    xs = gen.uexp(85, ncols, -30, 30)
    xs[::7] = 0.0
    x = o.set_d(xs, D)
    want = o.spmv(m, irp, ja, as_, x, 1, D)
    got = o.empty(m)
    counts = np.zeros(2, np.int32)
    assert emu.emu_spmv_split(sid, m, P(np.ascontiguousarray(irp, np.int32)), P(np.ascontiguousarray(ja, np.int32)),
                              P(np.ascontiguousarray(as_)), P(x), P(got), P(counts)) == 0
    assert got.tobytes() == want.tobytes(), name
    assert counts[0] > 0  # long rows reached their kernel


@pytest.mark.parametrize("name,lo,hi", [("16_4double", -30, 30), ("32_8double", -30, 30), ("32_8double", -400, 400),
                                        ("16_32bit", -40, 40)])
@pytest.mark.parametrize("n", [6, 7])
@pytest.mark.parametrize("ub", [0, 1])
def test_gemm_chain_nn(emu, name, lo, hi, n, ub):
    """The v1 gemm's chain kernel (SoA operands at 32-bit offsets, B's fields wave-uniform
    when UB) against the generic chain kernel on the same operands: every entry's sign,
    exponent and evaluation, or its REDO mark, agree; zeros and exact cancellations abandon
    some chains"""
    o = po.Oracle(name)
    sid = po.set_names().index(name)
    m, k = 9, 21
    As, Bs = gen.uexp(91, m * k, lo, hi), gen.uexp(92, k * n, lo, hi)
    As[4] = 0.0
    Bs[::5] = 0.0
    # entry (2, 1): its first two products cancel exactly
    As[2 + m * 0], As[2 + m * 1] = 1.5, 1.5
    Bs[0 + k * 1], Bs[1 + k * 1] = 0.75, -0.75
    A, B = o.set_d(As, D), o.set_d(Bs, D)
    w1, w2 = o.empty(m * n), o.empty(m * n)
    assert emu.emu_gemm_chain_nn(sid, m, n, k, P(A), P(B), P(w1), P(w2), ub) == 0
    assert w1.tobytes() == w2.tobytes()
    redo = (w1["sign"] != 0) & (w1["sign"] != 1)
    assert redo.any() and (hi > 100 or not redo.all())  # wide exponents abandon every chain

