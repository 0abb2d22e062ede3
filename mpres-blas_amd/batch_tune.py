"""Autotuner of the batched form's geometry (mpres_batch_tuning_t).

The geometry of the persistent batch kernel - lanes per number, persistent
workgroups per compute unit, gemv output rows per work item - never changes a
result, only the time. tune() plans the same batch under every candidate
geometry, times each plan's graph on the GPU and returns the fastest; save()
writes the winner with the measured table to a JSON file, load() reads it
back as the tuning argument of mpres.Batch. It plays the part of the
reference's compile-time launch constants (src/kernel_config.cuh), chosen by
measurement instead of by hand.

    python batch_tune.py [--set 16_4double] [--problems 4096] [--out tuning/batch_gfx950.json]

tunes on the batching workload of bench.py (alternating dots of log-uniform
length in [64, 8192] and square gemvs of log-uniform size in [32, 1024]).
"""
import argparse
import itertools
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

WG_PER_CU = (1, 2, 4)
GEMV_ROWS = (16, 32, 64, 128)


def candidates(ctx, lanes=None, wg_per_cu=WG_PER_CU, gemv_rows=GEMV_ROWS):
    lanes = lanes or [L for L in (1, 2, 4, 8, 16, 32) if ctx.L.mpres_lanes_supported(ctx.h, L)]
    return [dict(lanes=L, wg_per_cu=w, gemv_rows=r) for L, w, r in itertools.product(lanes, wg_per_cu, gemv_rows)]


def time_plan(mp, ctx, problems, tuning, reps=3):
    """average milliseconds of one graph launch of the plan (one untimed launch first);
    mp: the mpres binding module the problems were built with"""
    b = mp.Batch(ctx, problems, tuning)
    try:
        b.run()
        mp.synchronize()
        t = mp.Timer()
        t.start()
        for _ in range(reps):
            b.run()
        t.stop()
        return t.ms() / reps
    finally:
        b.close()


def tune(mp, ctx, problems, cands=None, reps=3, log=None):
    """time every candidate geometry; returns (best tuning, [(tuning, ms), ...])"""
    cands = cands or candidates(ctx)
    table = []
    for cand in cands:
        ms = time_plan(mp, ctx, problems, cand, reps)
        table.append((cand, ms))
        if log:
            log(f"[tune] {cand} {ms:.3f} ms")
    best = min(table, key=lambda e: e[1])[0]
    return best, table


def save(path, set_name, workload, best, table):
    data = {}
    if os.path.exists(path):
        with open(path) as f:
            data = json.load(f)
    data[set_name] = {"workload": workload, "best": best,
                      "measured_ms": [{"tuning": t, "ms": round(ms, 4)} for t, ms in table],
                      "date": time.strftime("%Y-%m-%d")}
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(data, f, indent=1)


def load(path, set_name):
    """the stored best geometry for a moduli set, or None"""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    e = data.get(set_name)
    return dict(e["best"]) if e else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--set", default="16_4double")
    ap.add_argument("--problems", type=int, default=4096)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning", "batch_gfx950.json"))
    args = ap.parse_args()
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "tests"))
    import gen  # noqa: E402
    import mpres  # noqa: E402

    ctx = mpres.Context(args.set)
    wl = gen.batch_workload(ctx, mpres, args.problems)
    log = lambda s: print(s, flush=True)  # noqa: E731
    best, table = tune(mpres, ctx, wl["problems"], reps=args.reps, log=log)
    save(args.out, args.set, wl["name"], best, table)
    print(json.dumps({"best": best, "ms": min(ms for _, ms in table)}), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
