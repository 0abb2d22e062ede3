"""Autotuner of the lanes per number of single calls, by operation and size
(mpres_set_lanes_table).

The lanes per number of an operation's kernels never change a result, only the
time, and the fastest count depends on the call's size: few lanes per number
give more numbers per wave, many lanes shorten each number's critical path. For
each moduli set, operation and size of a fixed list this script times every
compiled lane count on synthetic inputs (u 2^e, e in [-20, 20]; gemv and gemm with
beta = 0, so that repeated calls do the same work) and keeps the fastest. The table it writes (one entry per size, the bound of each bucket
halfway, geometrically, to the next size) plays the part of the reference's
compile-time launch constants (src/kernel_config.cuh), chosen by measurement
per size; load() installs it in a context.

    python lanes_tune.py [--sets 8_2double,16_4double,32_8double] [--out tuning/lanes_gfx950.json]
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

# operation -> sizes tuned (the call's multiply-accumulate count; SpMV: rows)
SIZES = {
    "dot": [1 << 12, 1 << 16, 1 << 20, 1 << 24],
    "axpy": [1 << 12, 1 << 16, 1 << 20, 1 << 24],
    "gemv": [64 * 64, 256 * 256, 1024 * 1024, 4096 * 4096],
    "spmv": [10 ** 4, 10 ** 5, 10 ** 6],
    "gemm": [64 ** 3, 128 ** 3, 256 ** 3, 512 ** 3],
    "elem": [1 << 12, 1 << 16, 1 << 20],
}


def op_code(mp, op):
    return {"dot": mp.K_DOT, "axpy": mp.K_AXPY, "gemv": mp.K_GEMV, "spmv": mp.K_SPMV, "elem": mp.K_ELEM,
            "gemm": mp.K_GEMM}[op]


def bounds_of(sizes):
    """bucket bounds: geometric midpoints between consecutive sizes, the last size last"""
    return [int(math.sqrt(a * b)) for a, b in zip(sizes, sizes[1:])] + [sizes[-1]]


def _timed(mp, fn, reps):
    fn()
    mp.synchronize()
    t = mp.Timer()
    t.start()
    for _ in range(reps):
        fn()
    t.stop()
    return t.ms() / reps


def make_call(mp, gen, ctx, op, size, keep):
    """a closure that runs one call of op at that size on fresh synthetic inputs"""
    L = ctx.L
    if op in ("dot", "axpy"):
        x, y = ctx.array_set_d(gen.uexp(11, size, -20, 20)), ctx.array_set_d(gen.uexp(12, size, -20, 20))
        r, al = ctx.array(1), ctx.array_set_d(gen.uexp(13, 1, -20, 20))
        keep += [x, y, r, al]
        if op == "dot":
            return lambda: mp.check(L.mpres_dot(ctx.h, size, C.byref(x), 1, C.byref(y), 1, C.byref(r), 256, 64,
                                                None, 0, None))
        return lambda: mp.check(L.mpres_axpy(ctx.h, size, C.byref(al), C.byref(x), 1, C.byref(y), 1, None))
    if op == "gemv":
        m = int(math.isqrt(size))
        A = ctx.array_set_d(gen.uexp(21, m * m, -20, 20))
        x, y = ctx.array_set_d(gen.uexp(22, m, -20, 20)), ctx.array_set_d(gen.uexp(23, m, -20, 20))
        # beta = 0: every repetition computes the same y (an accumulating y would drift in
        # exponent and change the work from one call to the next)
        al, be = ctx.array_set_d(gen.uexp(24, 1, -20, 20)), ctx.array_set_d([0.0])
        keep += [A, x, y, al, be]
        return lambda: mp.check(L.mpres_gemv(ctx.h, mp.NO_TRANS, m, m, C.byref(al), C.byref(A), m, C.byref(x), 1,
                                             C.byref(be), C.byref(y), 1, 32, None, 0, None))
    if op == "gemm":
        m = round(size ** (1 / 3))
        A, B = ctx.array_set_d(gen.uexp(51, m * m, -20, 20)), ctx.array_set_d(gen.uexp(52, m * m, -20, 20))
        Cm = ctx.array_set_d(gen.uexp(53, m * m, -20, 20))
        al, be = ctx.array_set_d(gen.uexp(54, 1, -20, 20)), ctx.array_set_d([0.0])  # beta = 0, as for gemv
        keep += [A, B, Cm, al, be]
        return lambda: mp.check(L.mpres_gemm(ctx.h, mp.NO_TRANS, mp.NO_TRANS, m, m, m, C.byref(al), C.byref(A), m,
                                             C.byref(B), m, C.byref(be), C.byref(Cm), m, None, 0, None))
    if op == "spmv":
        irp, ja, as_ = gen.powerlaw_csr(31, size)
        xs, ys = ctx.set_d(gen.uexp(32, size, -10, 10)), ctx.aos(count=size)
        d = [mp.DevBuf(host=irp), mp.DevBuf(host=ja), mp.DevBuf(host=as_)]
        csr = mp.Csr(d[2].ptr.value, d[1].ptr.value, d[0].ptr.value)
        keep += [xs, ys, csr] + d
        return lambda: mp.check(L.mpres_spmv_csr(ctx.h, size, C.byref(csr), xs.ptr, ys.ptr, None))
    xs, ys, rs = ctx.set_d(gen.uexp(41, size, -20, 20)), ctx.set_d(gen.uexp(42, size, -20, 20)), ctx.aos(count=size)
    keep += [xs, ys, rs]
    return lambda: mp.check(L.mpres_elementwise(ctx.h, mp.OP_ADD, size, xs.ptr, ys.ptr, rs.ptr, None))


def tune_set(mp, gen, set_name, ops, reps=3, log=None):
    """{op: {"sizes", "bounds", "lanes", "measured_ms"}} for one moduli set"""
    ctx = mp.Context(set_name)
    lanes = [L for L in (1, 2, 4, 8, 16, 32) if ctx.L.mpres_lanes_supported(ctx.h, L)]
    out = {}
    try:
        for op in ops:
            code = op_code(mp, op)
            best, table = [], []
            for size in SIZES[op]:
                keep = []
                fn = make_call(mp, gen, ctx, op, size, keep)
                row = {}
                for L in lanes:
                    ctx.set_lanes(code, L)
                    row[L] = _timed(mp, fn, reps)
                    if log:
                        log(f"[lanes] {set_name} {op} size {size} lanes {L}: {row[L]:.4f} ms")
                ctx.set_lanes(code, 0)
                best.append(min(row, key=row.get))
                table.append({"size": size, "ms": {str(k): round(v, 4) for k, v in row.items()}})
                del keep
            out[op] = {"sizes": SIZES[op], "bounds": bounds_of(SIZES[op]), "lanes": best, "measured_ms": table}
    finally:
        ctx.close()
    return out


def save(path, data):
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(data, f, indent=1)


def load(mp, ctx, set_name, path=os.path.join(HERE, "tuning", "lanes_gfx950.json")):
    """install the stored tables of a moduli set in a context; returns the operations set"""
    if not os.path.exists(path):
        return []
    with open(path) as f:
        data = json.load(f).get(set_name, {})
    ops = data.get("ops", {})
    for op, t in ops.items():
        ctx.set_lanes_table(op_code(mp, op), list(zip(t["bounds"], t["lanes"])))
    return sorted(ops)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sets", default="8_2double,16_4double,32_8double")
    ap.add_argument("--ops", default=",".join(SIZES))
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning", "lanes_gfx950.json"))
    args = ap.parse_args()
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "tests"))
    import gen  # noqa: E402
    import mpres  # noqa: E402

    log = lambda s: print(s, flush=True)  # noqa: E731
    data = {}
    for s in args.sets.split(","):
        data[s] = {"ops": tune_set(mpres, gen, s, args.ops.split(","), args.reps, log), "date": time.strftime("%Y-%m-%d")}
        save(args.out, data)
    print(json.dumps({s: {op: t["lanes"] for op, t in d["ops"].items()} for s, d in data.items()}), flush=True)


if __name__ == "__main__":
    main()
