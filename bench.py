#!/usr/bin/env python3
"""Throughput of the plain RNS multiple-precision path at 424 bits on one MI355X.

Workload (one step): mp_gemv y = alpha*A*x + beta*y with A 4096 x 4096, then
mp_gemm C = alpha*A*B + beta*C with 512 x 512 x 512, both with the 32-modulus
set 32_8double (424-bit precision). Inputs are u*2^e with u uniform in [-1,1)
and e uniform in [-30,30], converted to the mp format on the GPU before the
timed region. gemv uses the reference harness's row-sum block size 32, which
fixes its summation order.

Prints one JSON line: whole-step throughput in mp-FMA/s (one mp-FMA = one
multiply-accumulate of the operation: m*n for gemv, M*N*K for gemm), the
roofline position of the dominant kernel (its average duration measured with
HIP events around its launches), the relative error of sampled outputs
against the exact (dyadic) result, and the CPU reference algorithm timed on a
bounded sample of the same workload on this host.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time
from fractions import Fraction

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "tests")]

import gen  # noqa: E402

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E
PROF_GEMM_RESID = 6        # library timing slot of the gemm residue path (side stream)
PROF_GEMV_PAIR = 9         # library timing slot of the gemv residue and chain kernels (two streams)
# Constants from the committed counter profile of this workload (PROFILE below): HBM bytes
# per launch from rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE, each scaled by its
# reading for a 256 MiB device copy (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0), and the VALU
# wave-instructions per launch of the gemm chains (SQ_INSTS_VALU). They are not measured
# in this run, and the JSON line names them as profile values.
PROFILE = "profiles/bench_424bit_counters.md"
TRAFFIC_BYTES = {"chain": 148.2e6, "gemv": 2319.9e6, "gemv_chain": 685.8e6}
CHAIN_VALU_INSTS = 3.38e8
# VALU issue peak: 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction
VALU_ISSUE_PEAK = 614.4e9
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 (vector and matrix)
INT8_PEAK_TOPS = 5000.0    # MI355X dense INT8 MFMA

SET = "32_8double"
M_GEMV = N_GEMV = 4096
M_GEMM = N_GEMM = K_GEMM = 512
B3 = 32
ELO, EHI = -30, 30


def load_mpres():
    import importlib.util

    spec = importlib.util.spec_from_file_location("mpres", os.path.join(ROOT, "mpres-blas_amd", "mpres.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def exact_value(rec, m):
    M = 1
    for v in m:
        M *= int(v)
    X = 0
    for d, mi in zip(rec["digits"], m):
        mi = int(mi)
        Mi = M // mi
        X += int(d) * Mi * pow(Mi, -1, mi)
    X %= M
    v = Fraction(X) * (Fraction(2) ** int(rec["exp"]))
    return -v if int(rec["sign"]) else v


def cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable}


def _median_runs(fn, reps):
    """one untimed run, then the median wall time of reps runs"""
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_baseline(threads, reps=5):
    """The reference's CPU mp_float_t algorithm (the oracle's policy-E restatement), composed
    in the GPU's order, on this host. All threads: the full step (gemv 4096 x 4096 + gemm
    512^3, same input distribution), OpenMP over rows / entries. One thread: a row sample of
    the same step (gemv rows 0..255, gemm rows 0..31 of the same matrices; each output's
    arithmetic is that of the full problem). Each figure is the median of `reps` runs after
    one untimed run."""
    sys.path[:0] = [os.path.join(ROOT, "oracle")]
    import pyoracle as po

    o = po.Oracle(SET)
    E = po.POLICY_E
    Ad = gen.uexp(301, M_GEMV * N_GEMV, ELO, EHI)
    A = o.set_d(Ad, E)
    x = o.set_d(gen.uexp(302, N_GEMV, ELO, EHI), E)
    yd = gen.uexp(303, M_GEMV, ELO, EHI)
    y = o.set_d(yd, E)
    al = o.set_d(gen.uexp(304, 1, ELO, EHI), E)
    be = o.set_d(gen.uexp(305, 1, ELO, EHI), E)
    Agd = gen.uexp(311, M_GEMM * K_GEMM, ELO, EHI)
    Ag = o.set_d(Agd, E)
    Bg = o.set_d(gen.uexp(312, K_GEMM * N_GEMM, ELO, EHI), E)
    Cgd = gen.uexp(313, M_GEMM * N_GEMM, ELO, EHI)
    Cg = o.set_d(Cgd, E)

    def full():
        o.gemv(0, M_GEMV, N_GEMV, al, A, M_GEMV, x, be, y, B3, policy=E)
        o.gemm(M_GEMM, N_GEMM, K_GEMM, al, Ag, M_GEMM, Bg, K_GEMM, be, Cg, M_GEMM, policy=E)

    rv, rg = 256, 32  # sampled rows of the gemv and gemm outputs
    As = o.set_d(np.ascontiguousarray(Ad.reshape(N_GEMV, M_GEMV)[:, :rv]).ravel(), E)
    ys = o.set_d(yd[:rv], E)
    Ags = o.set_d(np.ascontiguousarray(Agd.reshape(K_GEMM, M_GEMM)[:, :rg]).ravel(), E)
    Cgs = o.set_d(np.ascontiguousarray(Cgd.reshape(N_GEMM, M_GEMM)[:, :rg]).ravel(), E)

    def sample():
        o.gemv(0, rv, N_GEMV, al, As, rv, x, be, ys, B3, policy=E)
        o.gemm(rg, N_GEMM, K_GEMM, al, Ags, rg, Bg, K_GEMM, be, Cgs, rg, policy=E)

    fmas = M_GEMV * N_GEMV + M_GEMM * N_GEMM * K_GEMM
    fmas_s = rv * N_GEMV + rg * N_GEMM * K_GEMM
    po.set_threads(threads)
    t_all, runs_all = _median_runs(full, reps)
    print(f"[bench] CPU {threads} threads: {t_all:.2f} s per step", file=sys.stderr, flush=True)
    po.set_threads(1)
    t_one, runs_one = _median_runs(sample, reps)
    print(f"[bench] CPU 1 thread: {t_one:.2f} s per sample", file=sys.stderr, flush=True)
    po.set_threads(threads)
    out = {
        "value": fmas / t_all,
        "unit": "mp-FMA/s",
        "cores": threads,
        "kind": "port",
        "sample": f"full step (gemv 4096x4096 + gemm 512^3, {fmas:.3g} mp-FMAs), median of {reps} runs after one "
                  f"untimed run, CPU mp_float_t algorithm (emulated directed rounding) composed in the GPU's order, "
                  f"OpenMP over rows/entries with {threads} threads",
        "runs_s": runs_all,
        "single_thread": {"value": fmas_s / t_one, "unit": "mp-FMA/s", "cores": 1,
                          "sample": f"gemv rows 0..{rv - 1} and gemm rows 0..{rg - 1} of the same step "
                                    f"({fmas_s:.3g} mp-FMAs), median of {reps} runs after one untimed run",
                          "runs_s": runs_one},
    }
    out.update(cpu_info())
    return out


def cpu_rate(fn, count, reps=5):
    """mp-FMA/s of the oracle's CPU algorithm for one call of fn (median of reps runs)"""
    t, _ = _median_runs(fn, reps)
    return count / t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-mfma", action="store_true", help="skip the precision-form (INT8 MFMA) gemm lines")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-form line (16 moduli, 4096 problems)")
    ap.add_argument("--no-other-ops", action="store_true",
                    help="skip the dot / axpy (106-bit) and SpMV (16 moduli) lines")
    ap.add_argument("--counters", action="store_true",
                    help="for hardware-counter runs: one warm-up and one measured call of the step's gemv and gemm, "
                         "then two 256 MiB device copies (the calibration of the byte counters); no JSON line")
    ap.add_argument("--counters-ops", action="store_true",
                    help="for hardware-counter runs: dot, axpy, SpMV, precision-form gemm and batched form, one "
                         "warm-up and one measured call each, then the calibration copies; no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    mp = load_mpres()
    if args.counters_ops:
        counter_ops_run(mp)
        return
    ctx = mp.Context(SET)
    nmod = ctx.n
    S = ctx.size
    print(f"[bench] context {SET}: n={nmod}, precision={ctx.precision} bits", file=sys.stderr, flush=True)

    # ---- inputs resident in HBM (converted on the GPU) ----
    A = ctx.array_set_d(gen.uexp(101, M_GEMV * N_GEMV, ELO, EHI))
    x = ctx.array_set_d(gen.uexp(102, N_GEMV, ELO, EHI))
    y0 = ctx.array_set_d(gen.uexp(103, M_GEMV, ELO, EHI))
    y = ctx.array_set_d(gen.uexp(103, M_GEMV, ELO, EHI))
    alpha = ctx.array_set_d(gen.uexp(104, 1, ELO, EHI))
    beta = ctx.array_set_d(gen.uexp(105, 1, ELO, EHI))
    Ag = ctx.array_set_d(gen.uexp(201, M_GEMM * K_GEMM, ELO, EHI))
    Bg = ctx.array_set_d(gen.uexp(202, K_GEMM * N_GEMM, ELO, EHI))
    C0 = ctx.array_set_d(gen.uexp(203, M_GEMM * N_GEMM, ELO, EHI))
    Cg = ctx.array_set_d(gen.uexp(203, M_GEMM * N_GEMM, ELO, EHI))
    alpha_g = ctx.array_set_d(gen.uexp(204, 1, ELO, EHI))
    beta_g = ctx.array_set_d(gen.uexp(205, 1, ELO, EHI))
    ws_bytes = ctx.L.mpres_gemv_workspace(ctx.h, mp.NO_TRANS, M_GEMV, N_GEMV, B3)
    ws = mp.DevBuf(nbytes=ws_bytes)
    mp.synchronize()
    print("[bench] inputs converted", file=sys.stderr, flush=True)

    L = ctx.L
    if args.counters:
        counter_run(ctx, mp, A, x, y, alpha, beta, Ag, Bg, Cg, alpha_g, beta_g, ws, ws_bytes)
        return
    ev = [mp.Timer() for _ in range(2)]
    t_gemv = [0.0]
    t_gemm = [0.0]

    def step(timed):
        if timed:
            ev[0].start()
        ctx.copy_array(y, y0, M_GEMV)
        mp.check(L.mpres_gemv(ctx.h, mp.NO_TRANS, M_GEMV, N_GEMV, C.byref(alpha), C.byref(A), M_GEMV, C.byref(x), 1,
                              C.byref(beta), C.byref(y), 1, B3, ws.ptr, ws_bytes, None))
        if timed:
            ev[0].stop()
            ev[1].start()
        ctx.copy_array(Cg, C0, M_GEMM * N_GEMM)
        mp.check(L.mpres_gemm(ctx.h, mp.NO_TRANS, mp.NO_TRANS, M_GEMM, N_GEMM, K_GEMM, C.byref(alpha_g), C.byref(Ag),
                              M_GEMM, C.byref(Bg), K_GEMM, C.byref(beta_g), C.byref(Cg), M_GEMM, None, 0, None))
        if timed:
            ev[1].stop()
            t_gemv[0] += ev[0].ms()
            t_gemm[0] += ev[1].ms()

    for _ in range(args.warmup):
        step(False)
    mp.synchronize()
    print("[bench] warmup done", file=sys.stderr, flush=True)

    # ---- timed region: K steps, no per-op events, kernel-level events inside the library ----
    ctx.profile(True)
    ctx.profile_read(mp.K_GEMM)
    ctx.profile_read(PROF_GEMM_RESID)
    ctx.profile_read(mp.K_GEMV)
    ctx.profile_read(PROF_GEMV_PAIR)
    mp.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(False)
    mp.synchronize()
    wall = time.perf_counter() - t0
    gemm_ms, gemm_n = ctx.profile_read(mp.K_GEMM)
    resid_ms, resid_n = ctx.profile_read(PROF_GEMM_RESID)
    gemv_ms, gemv_n = ctx.profile_read(mp.K_GEMV)
    pair_ms, pair_n = ctx.profile_read(PROF_GEMV_PAIR)
    ctx.profile(False)
    ms_per_step = wall * 1e3 / args.steps
    print(f"[bench] timed {args.steps} steps: {ms_per_step:.3f} ms/step", file=sys.stderr, flush=True)

    # per-operation split (separate pass with events between the operations)
    for _ in range(max(1, min(args.steps, 3))):
        step(True)
    nsplit = max(1, min(args.steps, 3))
    gemv_op_ms, gemm_op_ms = t_gemv[0] / nsplit, t_gemm[0] / nsplit

    fma_gemv = M_GEMV * N_GEMV
    fma_gemm = M_GEMM * N_GEMM * K_GEMM
    value = (fma_gemv + fma_gemm) / (ms_per_step / 1e3)

    # ---- roofline of the dominant kernel (the gemm chains or the gemv residue kernel) ----
    gemm_avg = gemm_ms / max(gemm_n, 1)
    resid_avg = resid_ms / max(resid_n, 1)
    gemv_avg = gemv_ms / max(gemv_n, 1)
    pair_avg = pair_ms / max(pair_n, 1)
    # the gemv's two streaming kernels run together: the residue kernel reads A's residues,
    # signs and exponents (4n + 8 bytes per element), the chain kernel A's signs and interval
    # bounds (40 bytes per element); both are HBM-bound
    gemv_bytes = M_GEMV * N_GEMV * (4 * nmod + 8 + 40)
    roof_gemv = {"kernel": "k_gemv_resid (residues of A, second stream) with k_gemv_chain (interval bounds of A)",
                 "bound": "hbm", "achieved": gemv_bytes / (pair_avg / 1e3) / 1e9, "peak": HBM_PEAK_GBS,
                 "unit": "GB/s", "frac": gemv_bytes / (pair_avg / 1e3) / 1e9 / HBM_PEAK_GBS,
                 "traffic": TRAFFIC_BYTES["gemv"] + TRAFFIC_BYTES["gemv_chain"], "traffic_source": PROFILE,
                 "avg_ms": pair_avg, "resid_kernel_ms": gemv_avg,
                 "work": "m*n*(4n+8+40) bytes (residues, sign, exponent and interval bounds of A)"}
    if gemm_avg >= pair_avg:
        # one interval step per mp-FMA: per bound one product (mul, div by the
        # eval unit) and one sum, i.e. 6 directed-rounded FP64 operations
        ach = 6.0 * fma_gemm / (gemm_avg / 1e3) / 1e12
        roof = {"kernel": "k_gemm_chain_nn (scalar interval chains of the gemm entries)", "bound": "fp64-valu",
                "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS,
                "traffic": TRAFFIC_BYTES["chain"], "traffic_source": PROFILE,
                "avg_ms": gemm_avg, "work": "M*N*K interval steps, 6 directed-rounded FP64 operations each",
                "binding_resource": "VALU issue: an interval step is about 160 VALU instructions (alignment, "
                                    "exponent and sign decisions, directed rounding) for its 6 FP64 operations",
                "valu_issue_frac": CHAIN_VALU_INSTS / (gemm_avg / 1e3) / VALU_ISSUE_PEAK,
                "valu_insts_source": PROFILE}
    else:
        roof = roof_gemv

    # ---- accuracy: sampled outputs against the exact dyadic result ----
    m_ = [int(v) for v in ctx_moduli(ctx)]
    err = accuracy(ctx, mp, A, x, y0, y, alpha, beta, Ag, Bg, C0, Cg, alpha_g, beta_g, m_)

    mfma = None
    if not args.no_mfma:
        mfma = mfma_lines(mp)
    others = None
    if not args.no_other_ops:
        others = plain_lines(mp, args.cpu_threads, cpu=not args.no_cpu)
    batched = None
    if not args.no_batch:
        batched = batch_lines(mp)

    cpu = None
    if not args.no_cpu:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_threads)

    out = {
        "metric": "mp-FMA/s at 106 & 424 bits: MFMA/batched vs plain HIP, at stated MPFR error",
        "value": value,
        "unit": "mp-FMA/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 residues + f64 interval",
        "data": "synthetic",
        "config": {"workload": "plain mp_gemv 4096x4096 + mp_gemm 512x512x512, 424-bit (32 moduli)",
                   "moduli_set": SET, "inputs": "u*2^e, u~U[-1,1), e~U[-30,30]", "gemv_blockDim3": B3,
                   "lanes_per_number": {"gemv": gemv_lanes(ctx, mp), "gemm": gemm_lanes(ctx, mp)}},
        "breakdown": {"gemv_ms": gemv_op_ms, "gemm_ms": gemm_op_ms,
                      "gemv_mpfma_per_s": fma_gemv / (gemv_op_ms / 1e3),
                      "gemm_mpfma_per_s": fma_gemm / (gemm_op_ms / 1e3),
                      "gemv_resid_kernel_ms": gemv_avg, "gemm_chain_kernel_ms": gemm_avg,
                      "gemm_resid_path_ms": resid_avg,
                      "gemm_resid_path": "exponent minima, int8 limb planes and INT8 MFMA residue products on the "
                                         "library's side stream, overlapped with the chains (wall time)"},
        "accuracy": err,
        "roofline": roof,
        "roofline_gemv": roof_gemv,
        "mfma_form": mfma,
        "other_ops": others,
        "batched_form": batched,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def mfma_lines(mp, size=2048, samples=((0, 0), (1000, 7), (2047, 2047), (5, 1500), (777, 333), (1500, 1999))):
    """Precision form of gemm on INT8 MFMA (mpres_gemm_mfma) beside the plain path on the
    same inputs: M = N = K = size at 8 and 32 moduli, exponents in [-8, 8] and [-60, 60].
    Errors: relative error of sampled entries of alpha*A*B + beta*C against the exact result."""
    out = []
    for name in ("8_2double", "32_8double"):
        ctx = mp.Context(name)
        m_ = [int(v) for v in ctx_moduli(ctx)]
        for lo, hi in ((-8, 8), (-60, 60)):
            print(f"[bench] mfma form {name} e[{lo},{hi}]", file=sys.stderr, flush=True)
            A = ctx.array_set_d(gen.uexp(401, size * size, lo, hi))
            B = ctx.array_set_d(gen.uexp(402, size * size, lo, hi))
            C0 = ctx.array_set_d(gen.uexp(403, size * size, lo, hi))
            Cm = ctx.array_set_d(gen.uexp(403, size * size, lo, hi))
            al, be = ctx.array_set_d(gen.uexp(404, 1, lo, hi)), ctx.array_set_d(gen.uexp(405, 1, lo, hi))
            L = ctx.L

            def run(fn):
                ctx.copy_array(Cm, C0, size * size)
                mp.check(fn(ctx.h, mp.NO_TRANS, mp.NO_TRANS, size, size, size, C.byref(al), C.byref(A), size,
                            C.byref(B), size, C.byref(be), C.byref(Cm), size, None, 0, None))

            tm = mp.Timer()
            run(L.mpres_gemm_mfma)
            reps = 3
            t_mf = 0.0
            for _ in range(reps):
                ctx.copy_array(Cm, C0, size * size)
                tm.start()
                mp.check(L.mpres_gemm_mfma(ctx.h, mp.NO_TRANS, mp.NO_TRANS, size, size, size, C.byref(al),
                                           C.byref(A), size, C.byref(B), size, C.byref(be), C.byref(Cm), size,
                                           None, 0, None))
                tm.stop()
                t_mf += tm.ms()
            t_mf /= reps
            err_mf = sample_error(ctx, mp, A, B, C0, Cm, al, be, size, samples, m_)
            # the product alone: alpha = 1, beta = 0 make the epilogue exact
            one, zero = ctx.array_set_d(np.array([1.0])), ctx.array_set_d(np.array([0.0]))
            ctx.copy_array(Cm, C0, size * size)
            mp.check(L.mpres_gemm_mfma(ctx.h, mp.NO_TRANS, mp.NO_TRANS, size, size, size, C.byref(one), C.byref(A),
                                       size, C.byref(B), size, C.byref(zero), C.byref(Cm), size, None, 0, None))
            err_prod = sample_error(ctx, mp, A, B, C0, Cm, one, zero, size, samples, m_)
            ctx.copy_array(Cm, C0, size * size)
            tm.start()
            mp.check(L.mpres_gemm(ctx.h, mp.NO_TRANS, mp.NO_TRANS, size, size, size, C.byref(al), C.byref(A), size,
                                  C.byref(B), size, C.byref(be), C.byref(Cm), size, None, 0, None))
            tm.stop()
            t_pl = tm.ms()
            err_pl = sample_error(ctx, mp, A, B, C0, Cm, al, be, size, samples, m_)
            fmas = float(size) ** 3
            out.append({"moduli_set": name, "bits": ctx.precision, "exponents": [lo, hi], "size": size,
                        "lanes_per_number": ctx.L.mpres_get_lanes_for(ctx.h, mp.K_MFMA, size ** 3),
                        "mfma_ms": t_mf, "mfma_mpfma_per_s": fmas / (t_mf / 1e3),
                        "plain_ms": t_pl, "plain_mpfma_per_s": fmas / (t_pl / 1e3), "speedup": t_pl / t_mf,
                        "mfma_max_rel_err": err_mf, "mfma_product_only_max_rel_err": err_prod,
                        "plain_max_rel_err": err_pl,
                        "unit_roundoff": 2.0 ** (-ctx.precision)})
            del A, B, C0, Cm
        ctx.close()
    return out


def time_call(mp, fn, reps):
    """average milliseconds of fn() over reps calls (one untimed call first), HIP events on the null stream"""
    fn()
    mp.synchronize()
    tm = mp.Timer()
    tm.start()
    for _ in range(reps):
        fn()
    tm.stop()
    return tm.ms() / reps


def exact_dot(a, b):
    """exact sum of a_i * b_i over float64 arrays, as a Fraction"""
    ma, ea = np.frexp(np.asarray(a, np.float64))
    mb, eb = np.frexp(np.asarray(b, np.float64))
    ia, ib = np.ldexp(ma, 53).astype(np.int64), np.ldexp(mb, 53).astype(np.int64)
    e = ea.astype(np.int64) + eb.astype(np.int64) - 106
    nz = (ia != 0) & (ib != 0)
    if not nz.any():
        return Fraction(0)
    ia, ib, e = ia[nz], ib[nz], e[nz]
    e0 = int(e.min())
    s = 0
    for p, q, k in zip(ia.tolist(), ib.tolist(), (e - e0).tolist()):
        s += (p * q) << k
    return Fraction(s) * Fraction(2) ** e0


def exact_dot_fast(a, b):
    """exact_dot for long vectors: each product p q 2^e (p, q integers below 2^53) is split
    into 16-bit pieces of its three 27-bit partial products (made nonnegative by a bias),
    the pieces are summed per exponent with np.bincount (sums below 2^40: exact in float64),
    and the per-exponent sums are recombined as Python integers"""
    ma, ea = np.frexp(np.asarray(a, np.float64))
    mb, eb = np.frexp(np.asarray(b, np.float64))
    ia, ib = np.ldexp(ma, 53).astype(np.int64), np.ldexp(mb, 53).astype(np.int64)
    e = ea.astype(np.int64) + eb.astype(np.int64) - 106
    nz = (ia != 0) & (ib != 0)
    if not nz.any():
        return Fraction(0)
    ia, ib, e = ia[nz], ib[nz], e[nz]
    e0 = int(e.min())
    k = e - e0
    ph, qh = ia >> 27, ib >> 27
    pl, ql = ia - (ph << 27), ib - (qh << 27)
    bias = 1 << 55
    nk = int(k.max()) + 1
    cnt = np.bincount(k, minlength=nk)
    total = 0
    for t, sh in ((ph * qh, 54), (ph * ql + pl * qh, 27), (pl * ql, 0)):
        u = t + bias  # 0 <= u < 2^56
        parts = [np.bincount(k, weights=((u >> (16 * j)) & 0xFFFF).astype(np.float64), minlength=nk)
                 for j in range(4)]
        for g in range(nk):
            if cnt[g]:
                sg = sum(int(parts[j][g]) << (16 * j) for j in range(4)) - bias * int(cnt[g])
                total += sg << (sh + g)
    return Fraction(total) * Fraction(2) ** e0


def rel_err(v, exact):
    if exact == 0:
        return 0.0 if v == 0 else float("inf")
    return float(abs(v - exact) / abs(exact))


def plain_lines(mp, threads, cpu=True):
    """The plain path's HBM-side operations: mp_dot and mp_axpy at 106 bits (8 moduli), the
    CSR SpMV at 16 moduli. Bytes are algorithmic (every operand read once, every result
    written once); errors are relative errors against the exact result of the fp64 inputs.
    Beside the dot and SpMV lines: the oracle's CPU algorithm (policy E) on the same inputs,
    with `threads` threads and with one."""
    sys.path[:0] = [os.path.join(ROOT, "oracle")]
    import pyoracle as po

    out = []
    ctx = mp.Context("8_2double")
    S = ctx.size
    m_ = [int(v) for v in ctx_moduli(ctx)]
    L = ctx.L
    for n, reps in ((65536, 20), (1 << 24, 5)):
        print(f"[bench] dot 106-bit n={n}", file=sys.stderr, flush=True)
        hx, hy = gen.uexp(1, n, -20, 20), gen.uexp(2, n, -20, 20)
        x, y, r = ctx.array_set_d(hx), ctx.array_set_d(hy), ctx.array(1)
        ms = time_call(mp, lambda: mp.check(L.mpres_dot(ctx.h, n, C.byref(x), 1, C.byref(y), 1, C.byref(r), 256, 64,
                                                        None, 0, None)), reps)
        if n <= 65536:
            want = exact_dot(hx, hy)
            if exact_dot_fast(hx, hy) != want:
                raise RuntimeError("exact_dot_fast disagrees with exact_dot")
        else:
            want = exact_dot_fast(hx, hy)
        err = rel_err(exact_value(ctx.array_to_host(r, 1)[0], m_), want)
        b = 2.0 * n * S
        line = {"op": "mp_dot", "moduli_set": "8_2double", "bits": ctx.precision, "n": n,
                "order": "gridDim3=256, blockDim3=64", "lanes_per_number": L.mpres_get_lanes_for(ctx.h, mp.K_DOT, n),
                "ms": ms, "mpfma_per_s": n / (ms / 1e3),
                "bytes": b, "gbs": b / (ms / 1e3) / 1e9, "hbm_frac": b / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
                "max_rel_err": err}
        if cpu and n <= 65536:
            o = po.Oracle("8_2double")
            ox, oy = o.set_d(hx, po.POLICY_E), o.set_d(hy, po.POLICY_E)
            fn = lambda: o.dot(ox, oy, 256, 64, po.POLICY_E)  # noqa: E731
            po.set_threads(threads)
            rt = cpu_rate(fn, n)
            po.set_threads(1)
            r1 = cpu_rate(fn, n)
            po.set_threads(threads)
            line["cpu"] = {"mpfma_per_s": rt, "threads": threads, "single_thread_mpfma_per_s": r1, "kind": "port",
                           "sample": "the same dot, CPU mp_float_t algorithm, median of 5 runs"}
        out.append(line)
        del x, y, r
    n = 1 << 24
    print(f"[bench] axpy 106-bit n={n}", file=sys.stderr, flush=True)
    x, y, al = ctx.array_set_d(gen.uexp(3, n, -20, 20)), ctx.array_set_d(gen.uexp(4, n, -20, 20)), ctx.array_set_d(
        gen.uexp(5, 1, -20, 20))
    ms = time_call(mp, lambda: mp.check(L.mpres_axpy(ctx.h, n, C.byref(al), C.byref(x), 1, C.byref(y), 1, None)), 5)
    b = 3.0 * n * S
    # accuracy: one call on a fresh y, sampled elements against the exact alpha*x_i + y_i
    hx, hy, ha = gen.uexp(3, n, -20, 20), gen.uexp(4, n, -20, 20), gen.uexp(5, 1, -20, 20)[0]
    y1 = ctx.array_set_d(hy)
    mp.check(L.mpres_axpy(ctx.h, n, C.byref(al), C.byref(x), 1, C.byref(y1), 1, None))
    mp.synchronize()
    aerr = 0.0
    for i in (0, 1, 4097, n // 2, n - 1):
        got = exact_value(fetch_range(ctx, mp, y1, i, 1)[0], m_)
        aerr = max(aerr, rel_err(got, Fraction(float(ha)) * Fraction(float(hx[i])) + Fraction(float(hy[i]))))
    out.append({"op": "mp_axpy", "moduli_set": "8_2double", "bits": ctx.precision, "n": n, "ms": ms,
                "lanes_per_number": L.mpres_get_lanes_for(ctx.h, mp.K_AXPY, n),
                "mpfma_per_s": n / (ms / 1e3), "bytes": b, "gbs": b / (ms / 1e3) / 1e9,
                "hbm_frac": b / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, "max_rel_err_sampled": aerr,
                "error_samples": "5 elements of y = alpha*x + y after one call, against the exact value"})
    del x, y, al, y1
    ctx.close()

    m = 1000000
    print(f"[bench] spmv 16 moduli, {m} rows (power-law)", file=sys.stderr, flush=True)
    irp, ja, as_ = gen.powerlaw_csr(7, m)
    nnz = int(irp[-1])
    ctx = mp.Context("16_4double")
    S = ctx.size
    m_ = [int(v) for v in ctx_moduli(ctx)]
    L = ctx.L
    hx = gen.uexp(8, m, -10, 10)
    xs, ys = ctx.set_d(hx), ctx.aos(count=m)
    dirp, dja, das = mp.DevBuf(host=irp), mp.DevBuf(host=ja), mp.DevBuf(host=as_)
    csr = mp.Csr(das.ptr.value, dja.ptr.value, dirp.ptr.value)
    rows = (0, 1, 12345, m - 1)
    want = {i: exact_dot(as_[irp[i]:irp[i + 1]], hx[ja[irp[i]:irp[i + 1]]]) for i in rows}
    b = nnz * 12.0 + (m + 1) * 4.0 + 2.0 * m * S
    for tpr in (None, 4, 8, 16, 32):
        if tpr is None:
            fn = lambda: mp.check(L.mpres_spmv_csr(ctx.h, m, C.byref(csr), xs.ptr, ys.ptr, None))  # noqa: E731
        else:
            fn = lambda: mp.check(L.mpres_spmv_csrv(ctx.h, m, C.byref(csr), tpr, xs.ptr, ys.ptr, None))  # noqa: E731
        ms = time_call(mp, fn, 5)
        yh = ctx.aos_to_host(ys, m)
        err = max(rel_err(exact_value(yh[i], m_), want[i]) for i in rows)
        out.append({"op": "mp_spmv_csr" if tpr is None else "mp_spmv_csrv", "threads_per_row": tpr,
                    "moduli_set": "16_4double", "bits": ctx.precision, "rows": m, "nnz": nnz, "ms": ms,
                    "lanes_per_number": L.mpres_get_lanes_for(ctx.h, mp.K_SPMV, m),
                    "mpfma_per_s": nnz / (ms / 1e3), "bytes": b, "gbs": b / (ms / 1e3) / 1e9,
                    "hbm_frac": b / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, "max_rel_err_sampled_rows": err})
    if cpu:
        o = po.Oracle("16_4double")
        ox = o.set_d(hx, po.POLICY_E)
        po.set_threads(threads)
        rt = cpu_rate(lambda: o.spmv(m, irp, ja, as_, ox, 1, po.POLICY_E), nnz)
        rows1 = np.arange(0, 100000, dtype=np.int32)
        nnz1 = int(irp[100000] - irp[0])
        po.set_threads(1)
        r1 = cpu_rate(lambda: o.spmv(m, irp, ja, as_, ox, 1, po.POLICY_E, rows=rows1), nnz1)
        po.set_threads(threads)
        out[-4]["cpu"] = {"mpfma_per_s": rt, "threads": threads, "single_thread_mpfma_per_s": r1, "kind": "port",
                          "sample": f"CSR SpMV, CPU mp_float_t algorithm: all {m} rows with {threads} threads, "
                                    f"rows 0..99999 ({nnz1} nonzeros) with one; medians of 5 runs"}
    ctx.close()
    return out


def batch_lines(mp, count=4096):
    """Batched form (mpres_batch_run) on the batching workload at 16 moduli: the geometry
    autotuned live (every candidate timed), the tuned plan's graph timed, and the same
    problems run as one library call per problem. Parity: sampled problems of the batch
    against their single calls, bit for bit; accuracy: sampled dots against the exact
    result of the fp64 inputs."""
    sys.path.insert(0, os.path.join(ROOT, "mpres-blas_amd"))
    import batch_tune

    ctx = mp.Context("16_4double")
    m_ = [int(v) for v in ctx_moduli(ctx)]
    L = ctx.L
    print(f"[bench] batch workload: {count} problems", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    wl = gen.batch_workload(ctx, mp, count)
    print(f"[bench] batch workload built in {time.perf_counter() - t0:.1f} s; tuning", file=sys.stderr, flush=True)
    probs = wl["problems"]
    best, table = batch_tune.tune(mp, ctx, probs, reps=2, log=lambda s: print(s, file=sys.stderr, flush=True))
    default_ms = batch_tune.time_plan(mp, ctx, probs, None, reps=2)
    plan = mp.Batch(ctx, probs, best)
    info = plan.info()
    ms = time_call(mp, plan.run, 5)

    def single(kind, a):
        if kind == "dot":
            mp.check(L.mpres_dot(ctx.h, a[0], C.byref(a[1]), a[2], C.byref(a[3]), a[4], C.byref(a[5]), a[6], a[7],
                                 None, 0, None))
        else:
            mp.check(L.mpres_gemv(ctx.h, a[0], a[1], a[2], C.byref(a[3]), C.byref(a[4]), a[5], C.byref(a[6]), a[7],
                                  C.byref(a[8]), C.byref(a[9]), a[10], a[11], None, 0, None))

    def singles():
        for kind, a in wl["singles"]:
            single(kind, a)

    print("[bench] batch: one call per problem", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ms_single = time_call(mp, singles, 2)
    wall_single = (time.perf_counter() - t0) / 3 * 1e3
    # parity on sampled problems: the batch's result against the single call from the same state
    sample = list(range(0, count, max(1, count // 16)))
    sample += [i + 1 for i in sample if i + 1 < count]
    saved = {}
    for i in sample:
        kind, a = wl["singles"][i]
        if kind == "gemv":
            saved[i] = ctx.array_to_host(a[9], a[1])
    plan.run()
    mp.synchronize()
    got = {}
    for i in sample:
        kind, a = wl["singles"][i]
        got[i] = ctx.array_to_host(a[5], 1) if kind == "dot" else ctx.array_to_host(a[9], a[1])
    mismatches = 0
    for i in sample:
        kind, a = wl["singles"][i]
        if kind == "gemv":
            mp.check(L.mpres_array_host2device(ctx.h, C.byref(a[9]), saved[i].ctypes.data_as(C.c_void_p), a[1], None))
        single(kind, a)
        mp.synchronize()
        ref = ctx.array_to_host(a[5], 1) if kind == "dot" else ctx.array_to_host(a[9], a[1])
        mismatches += int(ref.tobytes() != got[i].tobytes())
    errs = []
    for i in sample[:8]:
        kind, a = wl["singles"][i]
        if kind == "dot":
            s = 500 + 10 + 3 * i
            n = a[0]
            errs.append(rel_err(exact_value(got[i][0], m_), exact_dot(gen.uexp(s, n, -20, 20),
                                                                        gen.uexp(s + 1, n, -20, 20))))
    plan.close()
    fm, b = wl["fmas"], wl["bytes"]
    out = {"workload": wl["name"], "moduli_set": "16_4double", "bits": ctx.precision, "problems": count,
           "items": info["items"], "mp_fmas": fm, "batch_ms": ms, "batch_mpfma_per_s": fm / (ms / 1e3),
           "per_problem_ms": ms_single, "per_problem_wall_ms": wall_single,
           "per_problem_mpfma_per_s": fm / (ms_single / 1e3), "speedup": ms_single / ms,
           "bytes": b, "gbs": b / (ms / 1e3) / 1e9, "hbm_frac": b / (ms / 1e3) / 1e9 / HBM_PEAK_GBS,
           "tuning": best, "default_tuning_ms": default_ms, "candidates_timed": len(table),
           "sampled_problems_bit_identical_to_single_calls": f"{len(sample) - mismatches}/{len(sample)}",
           "sampled_dot_max_rel_err": max(errs) if errs else None}
    del wl
    ctx.close()
    return out


def sample_error(ctx, mp, A, B, C0, Cm, al, be, size, samples, m_):
    """max relative error of sampled entries of Cm against alpha*A*B + beta*C0 (exact)"""
    rows = sorted({i for i, _ in samples})
    Ah = dict(zip(rows, fetch_rows(ctx, mp, A, rows, size)))
    a = exact_value(ctx.array_to_host(al, 1)[0], m_)
    b = exact_value(ctx.array_to_host(be, 1)[0], m_)
    worst = 0.0
    for (i, j) in samples:
        colB = fetch_range(ctx, mp, B, j * size, size)
        ex = a * sum(exact_value(Ah[i][t], m_) * exact_value(colB[t], m_) for t in range(size))
        ex += b * exact_value(fetch_range(ctx, mp, C0, j * size + i, 1)[0], m_)
        got = exact_value(fetch_range(ctx, mp, Cm, j * size + i, 1)[0], m_)
        if ex != 0:
            worst = max(worst, float(abs(got - ex) / abs(ex)))
    return worst


def fetch_range(ctx, mp, X, start, count):
    """records start .. start+count-1 of an mp_array_t"""
    H = mp.hip()
    n = ctx.n
    dig = np.empty(count * n, np.int32)
    sg = np.empty(count, np.int32)
    ex = np.empty(count, np.int32)
    lo = np.empty(count, mp.ER_DTYPE)
    up = np.empty(count, mp.ER_DTYPE)
    length = array_len(mp, X)
    for dst, src, size in ((dig, X.digits + 4 * n * start, 4 * n * count), (sg, X.sign + 4 * start, 4 * count),
                           (ex, X.exp + 4 * start, 4 * count), (lo, X.eval + 16 * start, 16 * count),
                           (up, X.eval + 16 * (length + start), 16 * count)):
        if H.hipMemcpy(dst.ctypes.data_as(C.c_void_p), C.c_void_p(src), C.c_size_t(size), 2) != 0:
            raise RuntimeError("copy failed")
    recs = np.zeros(count, ctx.dtype)
    recs["digits"] = dig.reshape(count, n)
    recs["sign"] = sg
    recs["exp"] = ex
    recs["eval"][:, 0] = lo
    recs["eval"][:, 1] = up
    return recs


def counter_run(ctx, mp, A, x, y, alpha, beta, Ag, Bg, Cg, alpha_g, beta_g, ws, ws_bytes):
    L = ctx.L
    for _ in range(2):
        mp.check(L.mpres_gemv(ctx.h, mp.NO_TRANS, M_GEMV, N_GEMV, C.byref(alpha), C.byref(A), M_GEMV, C.byref(x), 1,
                              C.byref(beta), C.byref(y), 1, B3, ws.ptr, ws_bytes, None))
        mp.check(L.mpres_gemm(ctx.h, mp.NO_TRANS, mp.NO_TRANS, M_GEMM, N_GEMM, K_GEMM, C.byref(alpha_g), C.byref(Ag),
                              M_GEMM, C.byref(Bg), K_GEMM, C.byref(beta_g), C.byref(Cg), M_GEMM, None, 0, None))
    src, dst = mp.DevBuf(nbytes=256 << 20), mp.DevBuf(nbytes=256 << 20)
    H = mp.hip()
    for _ in range(2):
        if H.hipMemcpy(dst.ptr, src.ptr, C.c_size_t(256 << 20), 3) != 0:
            raise RuntimeError("device copy failed")
    mp.synchronize()
    print("[bench] counter run done", file=sys.stderr, flush=True)


def counter_ops_run(mp):
    """for hardware-counter runs of the other operations: one warm-up and one measured call of
    mp_dot and mp_axpy (n = 2^24, 8 moduli), the CSR and CSR-vector SpMV (10^6 rows, 16 moduli,
    4 threads per row), the precision-form gemm (2048^3, 8 moduli, e in [-8, 8]) and the
    batched form (1024 problems of the batching workload), then the 256 MiB calibration copies"""
    ctx = mp.Context("8_2double")
    L = ctx.L
    n = 1 << 24
    x, y, r = ctx.array_set_d(gen.uexp(1, n, -20, 20)), ctx.array_set_d(gen.uexp(2, n, -20, 20)), ctx.array(1)
    al = ctx.array_set_d(gen.uexp(5, 1, -20, 20))
    for _ in range(2):
        mp.check(L.mpres_dot(ctx.h, n, C.byref(x), 1, C.byref(y), 1, C.byref(r), 256, 64, None, 0, None))
    for _ in range(2):
        mp.check(L.mpres_axpy(ctx.h, n, C.byref(al), C.byref(x), 1, C.byref(y), 1, None))
    mp.synchronize()
    del x, y
    size = 2048
    A, B = ctx.array_set_d(gen.uexp(401, size * size, -8, 8)), ctx.array_set_d(gen.uexp(402, size * size, -8, 8))
    Cm = ctx.array_set_d(gen.uexp(403, size * size, -8, 8))
    be = ctx.array_set_d(gen.uexp(405, 1, -8, 8))
    for _ in range(2):
        mp.check(L.mpres_gemm_mfma(ctx.h, mp.NO_TRANS, mp.NO_TRANS, size, size, size, C.byref(al), C.byref(A), size,
                                   C.byref(B), size, C.byref(be), C.byref(Cm), size, None, 0, None))
    mp.synchronize()
    ctx.close()
    m = 1000000
    irp, ja, as_ = gen.powerlaw_csr(7, m)
    ctx = mp.Context("16_4double")
    L = ctx.L
    xs, ys = ctx.set_d(gen.uexp(8, m, -10, 10)), ctx.aos(count=m)
    dirp, dja, das = mp.DevBuf(host=irp), mp.DevBuf(host=ja), mp.DevBuf(host=as_)
    csr = mp.Csr(das.ptr.value, dja.ptr.value, dirp.ptr.value)
    for _ in range(2):
        mp.check(L.mpres_spmv_csr(ctx.h, m, C.byref(csr), xs.ptr, ys.ptr, None))
    for _ in range(2):
        mp.check(L.mpres_spmv_csrv(ctx.h, m, C.byref(csr), 4, xs.ptr, ys.ptr, None))
    mp.synchronize()
    wl = gen.batch_workload(ctx, mp, 1024)
    b = mp.Batch(ctx, wl["problems"], dict(lanes=4, wg_per_cu=4, gemv_rows=128))
    for _ in range(2):
        b.run()
    mp.synchronize()
    b.close()
    src, dst = mp.DevBuf(nbytes=256 << 20), mp.DevBuf(nbytes=256 << 20)
    H = mp.hip()
    for _ in range(2):
        if H.hipMemcpy(dst.ptr, src.ptr, C.c_size_t(256 << 20), 3) != 0:
            raise RuntimeError("device copy failed")
    mp.synchronize()
    ctx.close()
    print("[bench] counter run (other operations) done", file=sys.stderr, flush=True)


def ctx_bits(ctx):
    """bit length of the largest modulus"""
    return max(int(v) for v in ctx_moduli(ctx)).bit_length()


def ctx_moduli(ctx):
    out = np.zeros(ctx.n, np.int32)
    ctx.L.mpres_ctx_table(ctx.h, 0, out.ctypes.data_as(C.c_void_p), out.nbytes)
    return out


def gemv_lanes(ctx, mp):
    return ctx.L.mpres_get_lanes_for(ctx.h, mp.K_GEMV, M_GEMV * N_GEMV)


def gemm_lanes(ctx, mp):
    return ctx.L.mpres_get_lanes_for(ctx.h, mp.K_GEMM, M_GEMM * N_GEMM * K_GEMM)


def accuracy(ctx, mp, A, x, y0, y, alpha, beta, Ag, Bg, C0, Cg, al_g, be_g, m):
    """Relative error of 4 gemv outputs and 4 gemm entries vs the exact result of the same mp inputs."""
    prec = ctx.precision
    rows = [0, 1, 2047, 4095]
    yv = ctx.array_to_host(y, M_GEMV)
    xv = ctx.array_to_host(x, N_GEMV)
    y0v = ctx.array_to_host(y0, M_GEMV)
    a, b = exact_value(ctx.array_to_host(alpha, 1)[0], m), exact_value(ctx.array_to_host(beta, 1)[0], m)
    xs = [exact_value(r, m) for r in xv]
    worst = 0.0
    # rows of A: element (i, j) at j*4096 + i; fetch the needed rows through a small device gather
    Ah = fetch_rows(ctx, mp, A, rows, M_GEMV)
    for t, i in enumerate(rows):
        ex = b * exact_value(y0v[i], m) + a * sum(exact_value(Ah[t][j], m) * xs[j] for j in range(N_GEMV))
        got = exact_value(yv[i], m)
        if ex != 0:
            worst = max(worst, float(abs(got - ex) / abs(ex)))
    gv = worst
    worst = 0.0
    Cv, C0v = ctx.array_to_host(Cg, M_GEMM * N_GEMM), ctx.array_to_host(C0, M_GEMM * N_GEMM)
    Agv, Bgv = ctx.array_to_host(Ag, M_GEMM * K_GEMM), ctx.array_to_host(Bg, K_GEMM * N_GEMM)
    ag, bg = exact_value(ctx.array_to_host(al_g, 1)[0], m), exact_value(ctx.array_to_host(be_g, 1)[0], m)
    for (i, j) in [(0, 0), (1, 7), (300, 511), (511, 256)]:
        s = sum(exact_value(Agv[l * M_GEMM + i], m) * exact_value(Bgv[j * K_GEMM + l], m) for l in range(K_GEMM))
        ex = ag * s + bg * exact_value(C0v[j * M_GEMM + i], m)
        got = exact_value(Cv[j * M_GEMM + i], m)
        if ex != 0:
            worst = max(worst, float(abs(got - ex) / abs(ex)))
    return {"reference": "exact dyadic result of the same mp inputs (stronger than MPFR at 4x precision)",
            "gemv_max_rel_err": gv, "gemm_max_rel_err": worst, "unit_roundoff": 2.0 ** (-prec),
            "samples": "4 gemv outputs, 4 gemm entries"}


def array_len(mp, X):
    """the device-side length of an mp_array_t"""
    v = np.zeros(1, np.int32)
    if mp.hip().hipMemcpy(v.ctypes.data_as(C.c_void_p), C.c_void_p(X.len), C.c_size_t(4), 2) != 0:
        raise RuntimeError("copy failed")
    return int(v[0])


def fetch_rows(ctx, mp, A, rows, ld):
    """rows of a column-major mp_array_t with leading dimension and column count ld (host records)"""
    return [np.concatenate([fetch_range(ctx, mp, A, j * ld + i, 1) for j in range(ld)]) for i in rows]


if __name__ == "__main__":
    main()
