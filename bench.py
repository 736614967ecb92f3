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
# HBM bytes per launch of the FP32 path kernel on this workload, from the
# FETCH_SIZE and WRITE_SIZE counter passes, each scaled by the same counter's
# reading for a device copy of known size.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "k_paths_qe_fp32_4Mx365_hbm_traffic.txt")


# VALU issue utilisation of the path kernels from the SQ counter passes.
VALU_FILE = os.path.join(ROOT, "profiles", "k_paths_qe_4Mx365_valu_counters.md")
def _profile_value(path, key, conv):
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return conv(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def measured_traffic():
    return _profile_value(TRAFFIC_FILE, "traffic_bytes_per_launch", int)


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


def cpu_threads():
    """Threads for the CPU baseline: OMP_NUM_THREADS when the host sets it (the
    GPU box gives one GPU's process a share of 16 cores and sets it to 16),
    else every core this process may run on."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    return (int(env) if env.isdigit() and int(env) > 0 else avail), avail


def cpu_baseline(gpu32=None, gpu64=None, windows=5):
    """The FP64 CPU oracle (a port of the reference's per-path algorithm,
    OpenMP over paths) on the whole config-1 workload, split into `windows`
    equal path windows of the same random stream, each timed: the rate is the
    median window's.  The windows' moments merge into the full-size FP64
    oracle prices on the GPU's random numbers, against which the GPU's FP64
    and FP32 prices of the timed run are reported (relative and in SE)."""
    threads, avail = cpu_threads()
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    from oracle import oracle as o
    import quantkos_amd as qk
    c, pays = workload(qk.FP64)
    m = o.model(model=0, scheme=2, fp32_semantics=0, n_steps=N_STEPS, seed=42, S0=1.0, v0=0.04, r=0.03,
                q=0.0, T=1.0, kappa=2.0, theta=0.04, sigma=0.5, rho=-0.7, lambda_j=0.0, mu_j=0.0, sigma_j=0.0)
    op = [o.Payoff(p.type, p.right, p.K, p.B) for p in pays]
    w = N_PATHS // windows
    times, parts = [], []
    for k in range(windows):
        n_k = w if k < windows - 1 else N_PATHS - w * (windows - 1)
        t0 = time.perf_counter()
        parts.append(o.price_window(m, k * w, n_k, op, 100.0))
        times.append(time.perf_counter() - t0)
        progress("CPU baseline: window %d/%d, %d paths, %.2f s" % (k + 1, windows, n_k, times[-1]))
    med = sorted(times)[len(times) // 2]
    disc = math.exp(-0.03 * 1.0)
    full = [o.merge_stats([p[j] for p in parts], disc, 100.0) for j in range(len(op))]
    out = {"value": w * N_STEPS / med, "unit": "path-steps/s", "cores": threads, "kind": "port",
           "sample": "the whole workload (%d paths x %d steps, 3 payoffs, FP64) in %d windows of %d paths, "
                     "median window %.2f s (all: %s)" % (N_PATHS, N_STEPS, windows, w, med,
                                                          ", ".join("%.2f" % t for t in times)),
           "host_cores_available": avail, "nproc": os.cpu_count(),
           "oracle_full_size": [{"price": f.price, "se": f.se, "itm": f.itm} for f in full]}
    if gpu64 is not None:
        out["gpu_fp64_rel_err_vs_oracle"] = [abs(g.price - f.price) / abs(f.price) for g, f in zip(gpu64, full)]
        out["gpu_fp64_itm_equal"] = [g.itm == f.itm for g, f in zip(gpu64, full)]
    if gpu32 is not None:
        out["gpu_fp32_err_in_se_vs_oracle"] = [(g.price - f.price) / f.se for g, f in zip(gpu32, full)]
    return out


README_PUBLISHED_PRICE, README_PUBLISHED_SE = 5.110195, 0.003025
README_PUBLISHED_S = {"fp32": 0.0976, "fp64": 4.9191}   # README.md:334, :336 (RTX 3060 Laptop)


def other_configs(engine):
    """The other single-GPU workloads at their full sizes: Euler European FP64
    against the Heston characteristic-function price; Bates QE American put by
    Longstaff-Schwartz (HBM path store, per-date Gram reduction and solve); the
    Greek stencil as lanes of one launch; the reference's published README
    benchmark.  Each is timed once after a warm-up."""
    import quantkos_amd as qk
    out = {}
    # Heston European call, Euler full truncation, 1,000,000 paths x 252 steps.
    progress("Euler European 1M x 252 (FP64, FP32)")
    eu = {}
    for prec, name in ((qk.FP64, "fp64"), (qk.FP32, "fp32")):
        c = qk.contract(model=qk.HESTON, scheme=qk.EULER, precision=prec, n_steps=252, n_paths=1_000_000,
                        seed=42, S0=1.0, v0=0.04, r=0.03, q=0.0, T=1.0, kappa=2.0, theta=0.04, sigma=0.5,
                        rho=-0.7, price_scale=100.0)
        p = qk.payoff(qk.EUROPEAN, qk.CALL, 1.0)
        engine.price(c, p)
        t0 = time.perf_counter()
        st = engine.price(c, p)
        wall = time.perf_counter() - t0
        tm = engine.timing()
        cf = qk.analytic_european(c, qk.CALL, 1.0)      # the product's Heston characteristic-function price
        eu[name] = {"price": st.price, "se": st.se, "wall_ms": wall * 1e3, "kernel_ms": tm.path_kernel_ms,
                    "path_steps_per_s": 252e6 / wall, "err_in_se_vs_heston_cf": (st.price - cf) / st.se}
    eu["heston_cf"] = cf
    out["heston_euler_european_1Mx252"] = eu
    # Bates QE American put K=110, 500,000 paths, 50 exercise dates.
    progress("Bates QE American put 500k x 50 (FP32, FP64)")
    am = {}
    for prec, name, rs in ((qk.FP32, "fp32", 4), (qk.FP64, "fp64", 8)):
        c = qk.contract(model=qk.BATES, scheme=qk.QE, precision=prec, n_steps=50, n_paths=500_000, seed=42,
                        S0=1.0, v0=0.04, r=0.03, q=0.0, T=1.0, kappa=2.0, theta=0.04, sigma=0.5, rho=-0.7,
                        lambda_j=0.1, mu_j=-0.05, sigma_j=0.10, price_scale=100.0)
        engine.price_american(c, qk.PUT, 1.1)
        t0 = time.perf_counter()
        st = engine.price_american(c, qk.PUT, 1.1)
        wall = time.perf_counter() - t0
        tm = engine.timing()
        n, N = c.n_paths, c.n_steps
        # backward pass bytes: per path and date, read S(t) and S(t-1) and
        # read + write the cash flow
        bw_bytes = n * N * 4 * rs
        am[name] = {"price": st.price, "se": st.se, "wall_ms": wall * 1e3, "forward_ms": tm.path_kernel_ms,
                    "backward_ms": tm.lsm_backward_ms, "merge_ms": tm.reduce_ms,
                    "store_bytes": n * N * rs,
                    "forward_store_gb_s": n * N * rs / (tm.path_kernel_ms * 1e-3) / 1e9,
                    "backward_bytes": bw_bytes,
                    "backward_gb_s": bw_bytes / (tm.lsm_backward_ms * 1e-3) / 1e9,
                    "backward_us_per_date": tm.lsm_backward_ms * 1e3 / N}
    am["fp32_minus_fp64_in_se"] = (am["fp32"]["price"] - am["fp64"]["price"]) / am["fp64"]["se"]
    # the weighted Laguerre basis of degree 3 that the workload names
    # (exp(-X/2) L_k(X), Longstaff and Schwartz 2001; the reference fixes the
    # unweighted one at compile time, LSMAmerican.hpp:217-233)
    for prec, name in ((qk.FP32, "fp32"), (qk.FP64, "fp64")):
        c = qk.contract(model=qk.BATES, scheme=qk.QE, precision=prec, n_steps=50, n_paths=500_000, seed=42,
                        S0=1.0, v0=0.04, r=0.03, q=0.0, T=1.0, kappa=2.0, theta=0.04, sigma=0.5, rho=-0.7,
                        lambda_j=0.1, mu_j=-0.05, sigma_j=0.10, price_scale=100.0)
        engine.price_american(c, qk.PUT, 1.1, weighted=True)
        t0 = time.perf_counter()
        st = engine.price_american(c, qk.PUT, 1.1, weighted=True)
        wall = time.perf_counter() - t0
        tm = engine.timing()
        am["weighted_p03_" + name] = {"price": st.price, "se": st.se, "wall_ms": wall * 1e3,
                                      "forward_ms": tm.path_kernel_ms, "backward_ms": tm.lsm_backward_ms,
                                      "backward_us_per_date": tm.lsm_backward_ms * 1e3 / c.n_steps}
    out["bates_qe_american_put_500kx50"] = am
    # Greeks: Heston implicit Milstein, European call K=100 and cash-or-nothing
    # call K=105 from one path set, 2,000,000 paths x 252 steps, S0 +-1 %,
    # vol +-1e-3, r +-1bp, T -+1/365 (central), all bumps as lanes of one launch.
    progress("Greeks 2M x 252, 13 lanes x 2 contracts (FP32, FP64)")
    pays = [qk.payoff(qk.EUROPEAN, qk.CALL, 1.0), qk.payoff(qk.BINARY_CASH, qk.CALL, 1.05)]
    b = qk.bumps(rel_S=0.01, abs_vol=1e-3, abs_r=1e-4, abs_T=1.0 / 365.0, theta_central=1)
    names = ("delta", "gamma", "vega", "vomma", "vanna", "rho", "theta", "base_price", "base_se")
    gk = {}
    for prec, name in ((qk.FP32, "fp32"), (qk.FP64, "fp64")):
        c = qk.contract(model=qk.HESTON, scheme=qk.MILSTEIN, precision=prec, n_steps=252, n_paths=2_000_000,
                        seed=42, S0=1.0, v0=0.04, r=0.03, q=0.0, T=1.0, kappa=2.0, theta=0.04, sigma=0.5,
                        rho=-0.7, price_scale=100.0)
        engine.greeks(c, pays, b)
        t0 = time.perf_counter()
        g = engine.greeks(c, pays, b)
        wall = time.perf_counter() - t0
        tm = engine.timing()
        lanes = g[0].n_lanes
        gk[name] = {"lanes": lanes, "contracts": len(pays), "wall_ms": wall * 1e3, "kernel_ms": tm.path_kernel_ms,
                    "path_steps_per_s": lanes * c.n_paths * c.n_steps / wall,
                    "european_call": {k: getattr(g[0], k) for k in names},
                    "binary_cash_call": {k: getattr(g[1], k) for k in names}}
    gk["fp32_minus_fp64"] = {ct: {k: gk["fp32"][ct][k] - gk["fp64"][ct][k] for k in names[:-1]}
                             for ct in ("european_call", "binary_cash_call")}
    out["heston_milstein_greeks_2Mx252"] = gk
    # The reference's one published benchmark (README.md:324-345,
    # bench_asian_option_price_no_greeks.yaml): Heston QE arithmetic Asian call,
    # 5,000,000 paths x 365 steps, published 5.110195 (SE 0.003025) and
    # 0.0976 s (FP32 fast-math) / 4.9191 s (FP64) of "Total time spent for the
    # computations" on an RTX 3060 Laptop GPU.  wall_ms is the whole pricing
    # call (launch, merge, result copy), the counterpart of that line.
    progress("README benchmark: QE Asian 5M x 365 (FP32, FP64)")
    rd = {}
    for prec, name, pub_s in ((qk.FP32, "fp32", README_PUBLISHED_S["fp32"]), (qk.FP64, "fp64", README_PUBLISHED_S["fp64"])):
        c = qk.contract(model=qk.HESTON, scheme=qk.QE, precision=prec, n_steps=365, n_paths=5_000_000,
                        seed=88471920573105, S0=1.0, v0=0.05, r=0.03, q=0.02, T=1.0, kappa=5.0, theta=0.05,
                        sigma=0.5, rho=-0.8, price_scale=100.0)
        p = qk.payoff(qk.ASIAN, qk.CALL, 1.0)
        engine.price(c, p)
        t0 = time.perf_counter()
        st = engine.price(c, p)
        wall = time.perf_counter() - t0
        tm = engine.timing()
        rd[name] = {"price": st.price, "se": st.se, "wall_ms": wall * 1e3, "kernel_ms": tm.path_kernel_ms,
                    "path_steps_per_s": 5e6 * 365 / wall,
                    "err_vs_published_in_combined_se": (st.price - README_PUBLISHED_PRICE) /
                    math.sqrt(st.se ** 2 + README_PUBLISHED_SE ** 2),
                    "published_rtx3060_s": pub_s, "speedup_vs_published_other_hardware": pub_s / wall}
    out["readme_qe_asian_5Mx365"] = rd
    return out


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
    # the tuned schedule picks latency within 5 % of the best throughput; the
    # same schedule with the other number of batches in flight beside it
    tuned_inflight = int(((engine.tuning or {}).get("server") or {}).get("max_inflight", 2) or 2)
    other = 1 if tuned_inflight >= 2 else 2
    progress("serving: burst %.2f s; burst with %d batch(es) in flight" % (wall, other))
    with qk.Server(engine, max_inflight=other) as srv:
        t0 = time.perf_counter()
        srv.replay(cs, ps, [0.0] * len(cs))
        w2 = time.perf_counter() - t0
    out["burst"]["max_inflight"] = tuned_inflight
    out["burst_other_in_flight"] = {"max_inflight": other, "wall_s": w2, "path_steps_per_s": total_ps / w2}
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
    q = lambda v, f: v[min(len(v) - 1, int(f * len(v)))] * 1e3
    lat_am = sorted(r.latency_s for r, p in zip(res, ps2) if p.type == qk.AMERICAN)
    lat_rest = sorted(r.latency_s for r, p in zip(res, ps2) if p.type != qk.AMERICAN)
    out["poisson"] = {"rate_per_s": rate, "wall_s": wall, "path_steps_per_s": total_ps / wall,
                      "p50_ms": q(lat, 0.50), "p99_ms": q(lat, 0.99), "max_ms": lat[-1] * 1e3,
                      "american_p50_ms": q(lat_am, 0.50), "american_p99_ms": q(lat_am, 0.99),
                      "other_p50_ms": q(lat_rest, 0.50), "other_p99_ms": q(lat_rest, 0.99),
                      "p50_queue_ms": sorted(r.queue_s for r in res)[len(res) // 2] * 1e3,
                      "batches": st.n_batches, "failed": sum(r.status != 0 for r in res)}
    # 3) plain per-request FP64 launches (the drop-in single-contract path)
    progress("serving: Poisson %.2f s, p50 %.1f ms, p99 %.1f ms; plain FP64 launches"
             % (wall, out["poisson"]["p50_ms"], out["poisson"]["p99_ms"]))
    step = max(1, int(round(1.0 / fp64_fraction)))
    idx = list(range(0, len(cs), step))
    err, err_am = [], []
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
            (err_am if p.type == qk.AMERICAN else err).append((abs(burst[i].stats.price - s.price) / s.se, i))
    wall = time.perf_counter() - t0
    out["fp64_plain"] = {"requests": len(idx), "wall_s": wall, "path_steps_per_s": ps64 / wall}
    out["fp32_batched_over_fp64_plain"] = out["burst"]["path_steps_per_s"] / out["fp64_plain"]["path_steps_per_s"]
    def summary(e):
        e = sorted(e)
        if not e:
            return None
        return {"n": len(e), "max": e[-1][0], "max_request": e[-1][1], "p99": e[int(0.99 * (len(e) - 1))][0],
                "within_1se": sum(x <= 1.0 for x, _ in e) / len(e)}
    # FP32 batched vs FP64 plain, same random numbers.  For American requests
    # the two precisions may settle on different exercise policies, so their
    # difference also holds the policy's own sampling noise, which the
    # reported SE (cash flows under a fixed policy) does not measure.
    out["fp32_err_in_se_vs_fp64"] = summary(err)
    out["fp32_err_in_se_vs_fp64_american"] = summary(err_am)
    # Europeans against the semi-analytic price of their model (Heston or
    # Bates characteristic function): QE discretisation bias plus noise
    cf = []
    for i, (c, p) in enumerate(zip(cs, ps)):
        if p.type == qk.EUROPEAN and burst[i].stats.se > 0:
            cf.append((abs(burst[i].stats.price - qk.analytic_european(c, p.right, p.K)) / burst[i].stats.se, i))
    out["fp32_european_err_in_se_vs_cf"] = summary(cf)
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
    ap.add_argument("--no-other-configs", action="store_true")
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
    others = None if args.no_other_configs else other_configs(eng)
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
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "Heston Andersen-QE (psi_c=1.5, martingale correction), 4,000,000 paths x "
                               "365 steps, 3 payoffs from one path set: Asian call K=100, up-and-out "
                               "call K=100 B=130 daily, floating-strike lookback put",
                   "paths": N_PATHS, "time_steps": N_STEPS, "parallelism": "1 GPU"},
        "roofline": {"bound": "valu", "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / PEAK_FP32_TFLOPS, "traffic": measured_traffic(),
                     "kernel": "k_paths<float,Heston,QE,pathdep>", "kernel_ms": kavg32,
                     "valu_issue_util": _profile_value(VALU_FILE, "valu_issue_util_fp32", float),
                     "valu_instructions_per_path_step": _profile_value(
                         VALU_FILE, "valu_instructions_per_path_step_fp32", float),
                     "flops_per_path_step": FLOPS_PER_PS, "transcendentals_per_path_step": TRANS_PER_PS},
        "fp64": {"value": v64, "dtype": "f64", "ms_per_step": dt64 * 1e3 / s64, "kernel_ms": kavg64, "steps": s64,
                 "achieved_tflops": FLOPS_PER_PS * ps / (kavg64 * 1e-3) / 1e12,
                 "peak": PEAK_FP64_TFLOPS,
                 "valu_issue_util": _profile_value(VALU_FILE, "valu_issue_util_fp64", float)},
        "fp32_over_fp64": v32 / v64,
        "prices": {"fp32": [round(a.price, 6) for a in r32], "fp64": [round(b.price, 6) for b in r64],
                   "se_fp64": [round(b.se, 6) for b in r64], "fp32_err_in_se": [round(e, 4) for e in err_se]},
    }
    if others is not None:
        line["other_configs"] = others
    if serve is not None:
        line["serving"] = serve
    if not args.no_cpu_baseline:
        progress("CPU baseline (oracle, FP64)")
        line["cpu_baseline"] = cpu_baseline(r32, r64)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
