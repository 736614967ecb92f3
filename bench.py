"""Benchmark: path-steps/s of the fused Heston Andersen-QE path kernel, FP32 vs FP64.

Workload (one "step" = one full pricing): Heston QE (psi_c = 1.5, martingale
correction), 4,000,000 paths x 365 steps, S0=100, K=100, T=1, r=0.03, q=0,
v0=0.04, kappa=2, theta=0.04, sigma=0.5, rho=-0.7, seed 42, three payoffs from
one path set: arithmetic-average Asian call K=100, up-and-out call K=100 B=130
monitored daily, floating-strike lookback put.  Inputs are resident on the
device (the kernel generates its own random numbers); the timed region is K
back-to-back pricings bracketed by device synchronisations.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "path-steps/s Heston QE FP32 vs FP64 at ≤1 SE error; p50/p99 pricing latency"
N_PATHS, N_STEPS = 4_000_000, 365
# Algorithmic work per path-step of the QE kernel with three payoffs: 46 flops
# (quadratic branch incl. Box-Muller) + 3 payoff-tracking ops; 17 transcendentals
# (log, sqrt x4, sin, cos, exp, log, rcp x5 ...) counted separately.
FLOPS_PER_PS = 49
TRANS_PER_PS = 17
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6


def workload(precision, n_paths=N_PATHS):
    import quantkos_amd as qk
    c = qk.contract(model=qk.HESTON, scheme=qk.QE, precision=precision, n_steps=N_STEPS,
                    n_paths=n_paths, seed=42, S0=1.0, v0=0.04, r=0.03, q=0.0, T=1.0, kappa=2.0,
                    theta=0.04, sigma=0.5, rho=-0.7, price_scale=100.0)
    pays = [qk.payoff(qk.ASIAN, qk.CALL, 1.0), qk.payoff(qk.BARRIER_UP_OUT, qk.CALL, 1.0, 1.3),
            qk.payoff(qk.LOOKBACK_FLOATING, qk.PUT, 1.0)]
    return c, pays


def run(engine, precision, steps, warmup):
    c, pays = workload(precision)
    for _ in range(warmup):
        engine.price_payoffs(c, pays)
    engine.synchronize()
    kern = []
    t0 = time.perf_counter()
    for _ in range(steps):
        res = engine.price_payoffs(c, pays)
        kern.append(engine.timing().path_kernel_ms)
    engine.synchronize()
    dt = time.perf_counter() - t0
    return res, dt, kern


def cpu_baseline(target_s=15.0):
    """The FP64 CPU oracle (a port of the reference's per-path algorithm, OpenMP)
    on a bounded sample of the same workload, on this host's cores."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    os.environ["OMP_NUM_THREADS"] = str(cores)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    from oracle import oracle as o
    import quantkos_amd as qk
    c, pays = workload(qk.FP64)
    mk = lambda n: o.model(model=0, scheme=2, fp32_semantics=0, n_steps=N_STEPS, seed=42, S0=1.0,
                           v0=0.04, r=0.03, q=0.0, T=1.0, kappa=2.0, theta=0.04, sigma=0.5,
                           rho=-0.7, lambda_j=0.0, mu_j=0.0, sigma_j=0.0)
    op = [o.Payoff(p.type, p.right, p.K, p.B) for p in pays]
    n = 4096 * cores
    t0 = time.perf_counter()
    o.price(mk(n), n, op, 100.0)
    cal = time.perf_counter() - t0
    n2 = int(max(n, min(4_000_000, n * target_s / max(cal, 1e-3))))
    t0 = time.perf_counter()
    o.price(mk(n2), n2, op, 100.0)
    el = time.perf_counter() - t0
    return {"value": n2 * N_STEPS / el, "unit": "path-steps/s", "cores": cores, "kind": "port",
            "sample": "%d paths x %d steps of the same QE workload (3 payoffs), FP64, %.1f s" % (n2, N_STEPS, el)}


def progress(msg):
    print("[bench %s] %s" % (time.strftime("%H:%M:%S"), msg), file=sys.stderr, flush=True)


def serving(engine, n_requests=4096, fp64_fraction=0.25):
    """Config-4 request stream through the continuous-batching server (FP32),
    against plain per-request FP64 launches of the same requests."""
    import quantkos_amd as qk
    from quantkos_amd import workloads
    cs, ps, _ = workloads.config4_stream(n=n_requests, seed=2024, precision=qk.FP32)
    total_ps = workloads.path_steps(cs)
    mean_ps = total_ps / len(cs)
    out = {"requests": len(cs), "path_steps": total_ps, "scheme": "Andersen QE",
           "mix": "Euro 30/Asian 20/barrier 15/lookback 10/binary 5/American 20 %, Heston|Bates 50/50"}
    # 1) capacity: the whole stream submitted at t=0
    progress("serving: %d requests, %.3g path-steps, burst" % (len(cs), total_ps))
    with qk.Server(engine) as srv:
        t0 = time.perf_counter()
        burst = srv.replay(cs, ps, [0.0] * len(cs))
        wall = time.perf_counter() - t0
        st = srv.stats()
    cap = total_ps / wall
    out["burst"] = {"wall_s": wall, "path_steps_per_s": cap, "batches": st.n_batches,
                    "arena_peak_gb": st.arena_peak_bytes / 2 ** 30}
    # 2) Poisson arrivals at 0.8x the measured capacity
    rate = 0.8 * cap / mean_ps
    progress("serving: burst %.2f s, %.3g path-steps/s; Poisson at %.1f req/s" % (wall, cap, rate))
    cs2, ps2, arr = workloads.config4_stream(n=n_requests, seed=2024, precision=qk.FP32, rate=rate)
    with qk.Server(engine) as srv:
        t0 = time.perf_counter()
        res = srv.replay(cs2, ps2, arr)
        wall = time.perf_counter() - t0
        st = srv.stats()
    lat = sorted(r.latency_s for r in res)
    q = lambda f: lat[min(len(lat) - 1, int(f * len(lat)))]
    out["poisson"] = {"rate_per_s": rate, "wall_s": wall, "path_steps_per_s": total_ps / wall,
                      "p50_ms": q(0.50) * 1e3, "p99_ms": q(0.99) * 1e3, "max_ms": lat[-1] * 1e3,
                      "p50_queue_ms": sorted(r.queue_s for r in res)[len(res) // 2] * 1e3,
                      "batches": st.n_batches, "failed": sum(r.status != 0 for r in res)}
    # 3) plain per-request FP64 launches (the drop-in single-contract path)
    progress("serving: Poisson %.2f s, p50 %.1f ms, p99 %.1f ms; plain FP64 launches"
             % (wall, out["poisson"]["p50_ms"], out["poisson"]["p99_ms"]))
    step = max(1, int(round(1.0 / fp64_fraction)))
    idx = list(range(0, len(cs), step))
    err = []
    ps64 = 0
    t0 = time.perf_counter()
    for done, i in enumerate(idx):
        c = cs[i]
        c64 = qk.contract(**{f: getattr(c, f) for f, _ in c._fields_})
        c64.precision = qk.FP64
        p = ps[i]
        s = engine.price_american(c64, p.right, p.K) if p.type == qk.AMERICAN else engine.price(c64, p)
        ps64 += c.n_paths * c.n_steps
        if done % 128 == 0:
            progress("serving: plain FP64 %d/%d" % (done, len(idx)))
        if s.se > 0:
            err.append(abs(burst[i].stats.price - s.price) / s.se)
    wall = time.perf_counter() - t0
    out["fp64_plain"] = {"requests": len(idx), "wall_s": wall, "path_steps_per_s": ps64 / wall}
    out["fp32_batched_over_fp64_plain"] = out["burst"]["path_steps_per_s"] / out["fp64_plain"]["path_steps_per_s"]
    err.sort()
    out["fp32_err_in_se_vs_fp64"] = {"max": err[-1], "p99": err[int(0.99 * (len(err) - 1))],
                                     "within_1se": sum(e <= 1.0 for e in err) / len(err)}
    return out


def copy_calibration(n_bytes):
    """A plain device-to-device copy of known size (torch's copy kernel), run
    under `rocprofv3 --pmc FETCH_SIZE` / `WRITE_SIZE` to scale those counters
    to bytes.  Runs in its own process: torch ships its own HIP runtime."""
    import torch
    a = torch.ones(n_bytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n_bytes, "launches": 3}))


def main():
    if "--copy-calibration" in sys.argv:
        copy_calibration(int(sys.argv[sys.argv.index("--copy-calibration") + 1]))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serving-requests", type=int, default=4096)
    ap.add_argument("--serving-fp64-fraction", type=float, default=0.25)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import quantkos_amd as qk
    eng = qk.Engine(0)
    progress("config 1 FP32: %d warm-up + %d timed pricings" % (args.warmup, args.steps))
    r32, dt32, k32 = run(eng, qk.FP32, args.steps, args.warmup)
    progress("config 1 FP32: %.2f ms per pricing; FP64" % (dt32 * 1e3 / args.steps))
    s64 = args.fp64_steps or max(1, args.steps // 2)
    r64, dt64, k64 = run(eng, qk.FP64, s64, 1)
    serve = None if args.no_serving else serving(eng, args.serving_requests, args.serving_fp64_fraction)
    eng.close()
    ps = N_PATHS * N_STEPS
    v32 = ps * args.steps / dt32
    v64 = ps * s64 / dt64
    kavg32 = sum(k32) / len(k32)
    kavg64 = sum(k64) / len(k64)
    ach = FLOPS_PER_PS * ps / (kavg32 * 1e-3) / 1e12
    err_se = [abs(a.price - b.price) / b.se for a, b in zip(r32, r64)]
    line = {
        "metric": METRIC, "value": v32, "unit": "path-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt32 * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": "Heston Andersen-QE (psi_c=1.5, martingale correction), 4,000,000 paths x "
                               "365 steps, 3 payoffs from one path set: Asian call K=100, up-and-out "
                               "call K=100 B=130 daily, floating-strike lookback put",
                   "paths": N_PATHS, "time_steps": N_STEPS, "parallelism": "1 GPU"},
        "roofline": {"bound": "valu", "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "k_paths<float,Heston,QE,pathdep>", "kernel_ms": kavg32,
                     "flops_per_path_step": FLOPS_PER_PS, "transcendentals_per_path_step": TRANS_PER_PS},
        "fp64": {"value": v64, "ms_per_step": dt64 * 1e3 / s64, "kernel_ms": kavg64, "steps": s64,
                 "achieved_tflops": FLOPS_PER_PS * ps / (kavg64 * 1e-3) / 1e12,
                 "peak": PEAK_FP64_TFLOPS},
        "fp32_over_fp64": v32 / v64,
        "prices": {"fp32": [round(a.price, 6) for a in r32], "fp64": [round(b.price, 6) for b in r64],
                   "se_fp64": [round(b.se, 6) for b in r64], "fp32_err_in_se": [round(e, 4) for e in err_se]},
    }
    if serve is not None:
        line["serving"] = serve
    if not args.no_cpu_baseline:
        progress("CPU baseline (oracle, FP64)")
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
