"""Sparse formats on the CPU: the library's Matrix Market reader (mpres_mtx_*, host code)
against the file's entries, and the oracle's ELL / JAD / DIA SpMV pinned to its CSR SpMV
on matrices whose rows hold the same terms in the same order (so the sums must agree bit
for bit), for both policies."""
import ctypes as C
import importlib.util
import os

import numpy as np
import pytest

import gen
import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LF10 = os.path.join(ROOT, "tests", "golden", "LF10.mtx")


def _mp():
    spec = importlib.util.spec_from_file_location("mpres", os.path.join(ROOT, "mpres-blas_amd", "mpres.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


mp = _mp()


def P(a):
    return a.ctypes.data_as(C.c_void_p)


def file_entries(path):
    """(row, col, value) of the file's lines, 0-based, and whether it is symmetric"""
    with open(path) as f:
        sym = "symmetric" in f.readline()
        lines = [ln for ln in f if not ln.startswith("%")]
    m, n, nz = (int(v) for v in lines[0].split())
    ents = []
    for ln in lines[1:1 + nz]:
        i, j, v = ln.split()
        ents.append((int(i) - 1, int(j) - 1, float(v)))
    return m, n, ents, sym


def expanded(ents, sym):
    out = list(ents)
    if sym:
        out += [(j, i, v) for i, j, v in ents if i != j]
    return out


def write_general(path, seed=3, m=60, n=50):
    rng = np.random.default_rng(seed)
    ents = set()
    while len(ents) < 400:
        ents.add((int(rng.integers(0, m)), int(rng.integers(0, n))))
    ents = list(ents)
    rng.shuffle(ents)
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n% generated\n")
        f.write(f"{m} {n} {len(ents)}\n")
        for i, j in ents:
            f.write(f"{i + 1} {j + 1} {rng.uniform(-1, 1):.17g}\n")
    return path


def read_all(path):
    L = mp.lib()
    v = [C.c_int() for _ in range(6)]
    mp.check(L.mpres_mtx_info(path.encode(), *[C.byref(q) for q in v]))
    m, n, lines, maxnzr, nzmd, sym = (q.value for q in v)
    nnz = L.mpres_mtx_nnz(lines, nzmd, sym)
    csr = (np.zeros(nnz), np.zeros(nnz, np.int32), np.zeros(m + 1, np.int32))
    mp.check(L.mpres_mtx_csr(path.encode(), m, nnz, lines, sym, *[P(a) for a in csr]))
    ell = (np.zeros(m * maxnzr), np.zeros(m * maxnzr, np.int32))
    mp.check(L.mpres_mtx_ell(path.encode(), m, maxnzr, lines, sym, *[P(a) for a in ell]))
    jad = (np.zeros(nnz), np.zeros(nnz, np.int32), np.zeros(maxnzr + 1, np.int32), np.zeros(m, np.int32))
    mp.check(L.mpres_mtx_jad(path.encode(), m, maxnzr, nnz, lines, sym, *[P(a) for a in jad]))
    nd = C.c_int()
    mp.check(L.mpres_mtx_ndiag(path.encode(), lines, sym, C.byref(nd)))
    dia = (np.zeros(m * nd.value), np.zeros(nd.value, np.int32))
    mp.check(L.mpres_mtx_dia(path.encode(), m, lines, sym, nd.value, *[P(a) for a in dia]))
    dense = np.zeros(m * n)
    mp.check(L.mpres_mtx_dense(path.encode(), m, n, lines, sym, P(dense)))
    return dict(m=m, n=n, lines=lines, maxnzr=maxnzr, nzmd=nzmd, sym=sym, nnz=nnz, csr=csr, ell=ell, jad=jad,
                ndiag=nd.value, dia=dia, dense=dense.reshape(n, m).T)


@pytest.mark.parametrize("which", ["lf10", "general"])
def test_mtx_reader(which, tmp_path):
    path = LF10 if which == "lf10" else write_general(str(tmp_path / "g.mtx"))
    m, n, ents, sym = file_entries(path)
    ex = expanded(ents, sym)
    R = read_all(path)
    assert (R["m"], R["n"], R["lines"], R["sym"]) == (m, n, len(ents), int(sym))
    rows = [[] for _ in range(m)]
    for i, j, v in ex:
        rows[i].append((j, v))
    assert R["nnz"] == len(ex) and R["maxnzr"] == max(len(r) for r in rows)
    assert R["nzmd"] == sum(1 for i, j, _ in ents if i == j)
    # CSR: row pointers, and each row's (column, value) pairs as in the file
    as_, ja, irp = R["csr"]
    assert list(np.diff(irp)) == [len(r) for r in rows]
    for i in range(m):
        got = sorted(zip(ja[irp[i]:irp[i + 1]].tolist(), as_[irp[i]:irp[i + 1]].tolist()))
        assert got == sorted(rows[i])
    # ELL: slot c of row i is the c-th entry of row i in file order (mirrors at the line that makes them)
    eas, eja = R["ell"]
    cnt = [0] * m
    order = [[] for _ in range(m)]
    for i, j, v in ents:
        order[i].append((j, v))
        if sym and i != j:
            order[j].append((i, v))
    for i in range(m):
        for c in range(R["maxnzr"]):
            q = c * m + i
            if c < len(order[i]):
                assert (eja[q], eas[q]) == order[i][c]
            else:
                assert (eja[q], eas[q]) == (-1, 0.0)
        cnt[i] = len(order[i])
    # JAD: rows by decreasing length; diagonal j of permuted row r is entry j of CSR row perm[r]
    jas, jja, jcp, perm = R["jad"]
    lens = np.diff(irp)
    assert sorted(perm.tolist()) == list(range(m)) and all(np.diff(lens[perm]) <= 0)
    for j in range(R["maxnzr"]):
        assert jcp[j + 1] - jcp[j] == int((lens > j).sum())
        for r in range(jcp[j + 1] - jcp[j]):
            q, src = jcp[j] + r, irp[perm[r]] + j
            assert (jja[q], jas[q]) == (ja[src], as_[src])
    # DIA: sorted distinct offsets (mirrors appended), values on their diagonals; dense
    das, off = R["dia"]
    offs = sorted(set(j - i for i, j, _ in ents))
    if sym:
        offs += [-o for o in offs[-2::-1]]
    assert off.tolist() == offs
    D = np.zeros((m, n))
    for i, j, v in ex:
        D[i, j] = v
    assert np.array_equal(R["dense"], D)
    for d, o in enumerate(off):
        for i in range(m):
            j = i + o
            want = D[i, j] if 0 <= j < n else 0.0
            got = das[d * m + i]
            assert got == want or (got == 0.0 and want == 0.0)


@pytest.mark.parametrize("name", ["8_2double", "16_4double", "32_8double"])
@pytest.mark.parametrize("policy", [po.POLICY_D, po.POLICY_E])
def test_oracle_formats_match_csr(name, policy):
    """ELL, JAD (tpr 1 and 4) and DIA built from one CSR matrix, so each row holds the same
    terms in the same order: the oracle's four kernels give the CSR kernel's result"""
    o = po.Oracle(name)
    m = 700
    irp, ja, as_ = gen.powerlaw_csr(5, m, kmax=40)
    ja_sorted = ja.copy()
    for i in range(m):  # columns ascending within rows (the DIA order)
        s = np.argsort(ja[irp[i]:irp[i + 1]], kind="stable")
        ja_sorted[irp[i]:irp[i + 1]] = ja[irp[i]:irp[i + 1]][s]
        as_[irp[i]:irp[i + 1]] = as_[irp[i]:irp[i + 1]][s]
    ja = ja_sorted
    x = o.set_d(gen.uexp(6, m, -20, 20), policy)
    want = o.spmv(m, irp, ja, as_, x, 1, policy)
    lens = np.diff(irp)
    w = int(lens.max())
    # ELL
    eas, eja = np.zeros(m * w), np.full(m * w, -1, np.int32)
    for i in range(m):
        for c in range(lens[i]):
            eas[c * m + i], eja[c * m + i] = as_[irp[i] + c], ja[irp[i] + c]
    assert o.spmv_ell(m, w, eas, eja, x, policy).tobytes() == want.tobytes()
    # JAD
    perm = np.argsort(-lens, kind="stable").astype(np.int32)
    jcp = np.zeros(w + 1, np.int32)
    jas, jja = [], []
    for j in range(w):
        for r in perm:
            if lens[r] > j:
                jas.append(as_[irp[r] + j])
                jja.append(ja[irp[r] + j])
        jcp[j + 1] = len(jas)
    for tpr in (1, 4):
        got = o.spmv_jad(m, w, np.array(jas), np.array(jja, np.int32), jcp, perm, x, tpr, policy)
        assert got.tobytes() == o.spmv(m, irp, ja, as_, x, tpr, policy).tobytes()
    # DIA (ascending offsets = ascending columns in each row)
    offs = sorted(set((ja - np.repeat(np.arange(m), lens)).tolist()))
    das = np.zeros(m * len(offs))
    for i in range(m):
        for q in range(irp[i], irp[i + 1]):
            das[offs.index(int(ja[q]) - i) * m + i] = as_[q]
    got = o.spmv_dia(m, m, len(offs), das, np.array(offs, np.int32), x, policy)
    assert got.tobytes() == want.tobytes()
