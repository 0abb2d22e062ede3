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


def batch_workload(ctx, mp, count=4096, seed=500, elo=-20, ehi=20, dot_order=(1, 64), gemv_block3=8):
    """The batching configuration on the device: problem i even is a dot of integer
    length log-uniform in [64, 8192], i odd a square gemv y = alpha*A*x + beta*y of
    size log-uniform in [32, 1024]; values u*2^e with e uniform in [elo, ehi]; one
    alpha and one beta shared by all gemvs. Order arguments: dots grid3 x block3 =
    dot_order, gemvs block3 = gemv_block3. Returns the batch entries (mpres
    BatchProblem), the arrays they use, the single calls they stand for, the
    multiply-accumulate count and the algorithmic bytes (every operand read once,
    every result written once)."""
    rng = np.random.default_rng(seed)
    S = ctx.size
    keep, problems, singles = [], [], []
    fmas = nbytes = 0.0
    alpha = ctx.array_set_d(uexp(seed + 1, 1, elo, ehi))
    beta = ctx.array_set_d(uexp(seed + 2, 1, elo, ehi))
    keep += [alpha, beta]
    for i in range(count):
        s = int(seed + 10 + 3 * i)
        if i % 2 == 0:
            n = int(np.rint(np.exp(rng.uniform(np.log(64), np.log(8192)))))
            x, y, r = ctx.array_set_d(uexp(s, n, elo, ehi)), ctx.array_set_d(uexp(s + 1, n, elo, ehi)), ctx.array(1)
            keep += [x, y, r]
            problems.append(mp.dot_problem(n, x, y, r, *dot_order))
            singles.append(("dot", (n, x, 1, y, 1, r, dot_order[0], dot_order[1])))
            fmas += n
            nbytes += (2 * n + 1) * S
        else:
            m = int(np.rint(np.exp(rng.uniform(np.log(32), np.log(1024)))))
            A = ctx.array_set_d(uexp(s, m * m, elo, ehi))
            x, y = ctx.array_set_d(uexp(s + 1, m, elo, ehi)), ctx.array_set_d(uexp(s + 2, m, elo, ehi))
            keep += [A, x, y]
            problems.append(mp.gemv_problem(mp.NO_TRANS, m, m, alpha, A, m, x, beta, y, gemv_block3))
            singles.append(("gemv", (mp.NO_TRANS, m, m, alpha, A, m, x, 1, beta, y, 1, gemv_block3)))
            fmas += m * m
            nbytes += (m * m + 3 * m) * S
    name = (f"{count} problems ({ctx.set_name}): dots n~logU[64,8192] order {dot_order[0]}x{dot_order[1]}, "
            f"square gemvs m~logU[32,1024] block3 {gemv_block3}, e~U[{elo},{ehi}]")
    return {"problems": problems, "keep": keep, "singles": singles, "fmas": fmas, "bytes": nbytes, "name": name}
