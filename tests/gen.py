"""Seeded synthetic inputs of the measured configurations: x = u * 2^e with u
uniform in [-1, 1) (53-bit resolution) and an integer exponent e uniform in
[elo, ehi]; and a power-law CSR matrix for the SpMV configuration."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def uexp(seed, count, elo, ehi):
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 2**53, count, dtype=np.int64)
    u = np.ldexp(k.astype(np.float64), -52) - 1.0
    e = rng.integers(elo, ehi + 1, count)
    return np.ldexp(u, e)


def powerlaw_csr(seed, m, kmax=1024, alpha=1.6138):
    """Row lengths k_i ~ P(k) ~ k^-alpha on [1, kmax] (k_i counts the diagonal;
    alpha = 1.6138 gives a mean of 16); each row holds its diagonal plus
    k_i - 1 distinct uniform columns, sorted ascending; values uniform in [-1, 1)."""
    rng = np.random.default_rng(seed)
    ks = np.arange(1, kmax + 1, dtype=np.float64)
    p = ks ** (-alpha)
    p /= p.sum()
    k = rng.choice(np.arange(1, kmax + 1), size=m, p=p)
    k = np.minimum(k, m)
    irp = np.zeros(m + 1, np.int64)
    irp[1:] = np.cumsum(k)
    nnz = int(irp[-1])
    rows = np.repeat(np.arange(m, dtype=np.int64), k)
    ja = rng.integers(0, m, nnz, dtype=np.int64)
    ja[irp[:-1]] = np.arange(m)  # the diagonal
    # sort within rows; redraw duplicates (and off-diagonal hits of the diagonal)
    order = np.lexsort((ja, rows))
    ja = ja[order]
    dup = np.zeros(nnz, bool)
    dup[1:] = (ja[1:] == ja[:-1]) & (rows[1:] == rows[:-1])
    for i in np.unique(rows[dup]):
        a, b = irp[i], irp[i + 1]
        cols = {int(i)}
        while len(cols) < b - a:
            cols.add(int(rng.integers(0, m)))
        ja[a:b] = np.array(sorted(cols))
    as_ = rng.uniform(-1.0, 1.0, nnz)
    return irp.astype(np.int32), ja.astype(np.int32), as_


def read_mtx(path):
    """Matrix Market coordinate file -> CSR (symmetric files expanded), rows sorted by column."""
    with open(path) as f:
        lines = [ln for ln in f if not ln.startswith("%")]
    sym = "symmetric" in open(path).readline()
    m, n, nz = (int(v) for v in lines[0].split())
    ents = []
    for ln in lines[1: 1 + nz]:
        i, j, v = ln.split()
        i, j, v = int(i) - 1, int(j) - 1, float(v)
        ents.append((i, j, v))
        if sym and i != j:
            ents.append((j, i, v))
    ents.sort()
    irp = np.zeros(m + 1, np.int32)
    for i, _, _ in ents:
        irp[i + 1] += 1
    irp = np.cumsum(irp).astype(np.int32)
    ja = np.array([j for _, j, _ in ents], np.int32)
    as_ = np.array([v for _, _, v in ents], np.float64)
    return m, n, irp, ja, as_
