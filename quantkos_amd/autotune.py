"""Autotuner for the launch schedules of the path kernels.

Measures candidate schedules on the device and writes a JSON table that
`Engine` (lanes per thread, through qk_set_schedule) and `Server` (tile
schedule of the batch kernel) apply by default.  Every schedule gives the
same prices up to rounding, so tuning changes speed only.

Knobs:
  * lane schedule: how many lanes of a qk_price_lanes / qk_price_greeks launch
    one thread carries (1, 2 or 4, or 8 in FP32), keyed by (precision, scheme,
    n_lanes).
    More lanes per thread share one Philox block and one Box-Muller transform
    per step; fewer keep more waves resident and waste no padded lanes.
  * tile schedule of the request server: path-steps per work item (the tiles a
    wavefront pulls at a time), resident blocks per CU of a batch launch,
    batch launches in flight, and the work cap of a batch (items per resident
    wavefront).  Chosen by the burst throughput of the whole 4,096-request
    stream (its American requests cycle the server's arena several times,
    which a short stream does not show), then by p99 latency (p50 breaking
    ties) under Poisson arrivals among the candidates within 5 % of the best
    throughput.
    Fewer resident batch blocks leave registers for the Longstaff-Schwartz
    pump's blocks beside a batch; smaller batches end more often, so
    admission keeps the pump supplied.

The reference has no counterpart: its launch shape is whatever Kokkos'
RangePolicy picks (MSolver.hpp:195-232).

    python -m quantkos_amd.autotune [--quick] [--server-only] [--out FILE]
"""
from __future__ import annotations

import argparse
import json
import os
import time

from . import _lib

PKG = os.path.dirname(os.path.abspath(__file__))
TUNING_DIR = os.path.join(PKG, "tuning")
ARCH = os.environ.get("QK_OFFLOAD_ARCH", "gfx950")
DEFAULT_PATH = os.path.join(TUNING_DIR, ARCH + ".json")

LANES_PER_THREAD = (1, 2, 4, 8)      # 8 in FP32 only
# Lane counts tuned: the Greek stencils (12 lanes, 13 with the central theta
# lane, 15 with every bump) and generic scenario fans.
LANE_COUNTS = (2, 4, 8, 12, 13, 15, 16)
# Two batches in flight fill each other's tails but share the GPU; with a
# backlog gate of 2 the second one runs only while the queue holds two
# batches' work (bursts), so light traffic keeps one batch at a time.
SERVER_CANDIDATES = [dict(item_path_steps=65536, blocks_per_cu=bpc, max_inflight=mi, batch_items_per_wave=ipw,
                          overlap_min_backlog=gate)
                     for bpc in (3, 4) for mi, gate in ((1, 0), (2, 0), (2, 2)) for ipw in (16, 32)]
# Burst-throughput band within which the server schedule is chosen by latency:
# repeated burst measurements of one schedule differ by up to ~3 %.
SERVER_RATE_TOLERANCE = 0.05
# p99 band within which throughput decides: repeated Poisson replays of one
# schedule give p99 values a few per cent apart.
SERVER_P99_TOLERANCE = 0.05
SERVER_REPEATS = 2


def load(path: str = DEFAULT_PATH):
    """The table at `path`, or None when there is none."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    validate(table)
    return table


def validate(table: dict):
    for e in table.get("lanes", []):
        if e["lanes_per_thread"] not in LANES_PER_THREAD or (e["lanes_per_thread"] == 8 and e["precision"] != _lib.FP32):
            raise ValueError("lanes_per_thread must be 1, 2 or 4, or 8 in FP32: %r" % (e,))
        if not 1 <= e["n_lanes"] <= 64:
            raise ValueError("n_lanes must be in [1, 64]: %r" % (e,))
    srv = table.get("server")
    if srv is not None:
        for k in ("item_path_steps", "blocks_per_cu", "max_inflight", "batch_items_per_wave",
                  "overlap_min_backlog"):
            if int(srv.get(k, 0)) < 0:
                raise ValueError("server.%s must be >= 0" % k)


def apply(engine, table: dict):
    """Installs the lane schedule on the engine's ctx; the server defaults are
    read from engine.tuning when a Server is created."""
    validate(table)
    engine.set_schedule((e["precision"], e["scheme"], e["n_lanes"], e["lanes_per_thread"])
                        for e in table.get("lanes", []))
    engine.tuning = table


def pick_best(timings: dict):
    """Candidate with the smallest time; ties go to the first listed."""
    best = None
    for cand, t in timings.items():
        if best is None or t < timings[best]:
            best = cand
    return best


def _lane_contracts(precision, scheme, n_lanes, n_paths, n_steps):
    from .engine import contract
    base = dict(model=_lib.HESTON, scheme=scheme, precision=precision, n_steps=n_steps, n_paths=n_paths,
                seed=7, S0=1.0, v0=0.04, r=0.03, q=0.0, T=1.0, kappa=2.0, theta=0.04, sigma=0.5, rho=-0.7,
                price_scale=100.0)
    lanes = []
    for i in range(n_lanes):
        kw = dict(base)
        kw["S0"] = 1.0 + 0.01 * ((i % 5) - 2)
        kw["v0"] = 0.04 + 0.001 * (i // 5)
        lanes.append(contract(**kw))
    return lanes


def tune_lanes(engine, precisions=(_lib.FP32, _lib.FP64), schemes=(_lib.EULER, _lib.MILSTEIN, _lib.QE),
               lane_counts=LANE_COUNTS, n_paths=1 << 20, n_steps=64, repeats=3, log=print):
    """Times qk_price_lanes (European + Asian payoffs) for each lanes-per-thread
    candidate and returns the table entries with their kernel times."""
    from .engine import payoff
    pays = [payoff(_lib.EUROPEAN, _lib.CALL, 1.0), payoff(_lib.ASIAN, _lib.CALL, 1.0)]
    entries = []
    for prec in precisions:
        for scheme in schemes:
            for n in lane_counts:
                lanes = _lane_contracts(prec, scheme, n, n_paths, n_steps)
                times = {}
                for lpt in (l for l in LANES_PER_THREAD if l < 8 or prec == _lib.FP32):
                    engine.set_schedule([(prec, scheme, n, lpt)])
                    engine.price_lanes(lanes, pays)                  # warm-up
                    best = float("inf")
                    for _ in range(repeats):
                        engine.price_lanes(lanes, pays)
                        best = min(best, engine.timing().path_kernel_ms)
                    times[lpt] = best
                lpt = pick_best(times)
                entries.append({"precision": prec, "scheme": scheme, "n_lanes": n, "lanes_per_thread": lpt,
                                "kernel_ms": {str(k): round(v, 4) for k, v in times.items()}})
                log("lanes prec=%d scheme=%d n_lanes=%d -> %d per thread  %s"
                    % (prec, scheme, n, lpt, " ".join("%d:%.3fms" % kv for kv in times.items())))
    # leave the engine's own schedule as it was
    if engine.tuning:
        apply(engine, engine.tuning)
    else:
        engine.set_schedule([])
    return entries


def pick_server(rates: dict, latencies: dict, tolerance: float = SERVER_RATE_TOLERANCE,
                p99_tolerance: float = SERVER_P99_TOLERANCE):
    """Server schedule choice: of the candidates whose burst throughput is
    within `tolerance` of the best, those whose p99 latency under Poisson
    arrivals (`latencies` maps a candidate to (p50, p99)) is within
    `p99_tolerance` of the smallest, and of these the one with the highest
    throughput; ties go to the first listed.  A serving target is set on the
    tail: a schedule that lowers p50 by letting the tail grow (two batches in
    flight without the backlog gate starve the Longstaff-Schwartz pump of
    registers) loses, and p99 differences within a replay's spread do not
    decide.  Candidates without a latency entry count only when none has one."""
    top = max(rates.values())
    short = [k for k, v in rates.items() if v >= (1.0 - tolerance) * top]
    timed = [k for k in short if k in latencies]
    if not timed:
        return pick_best({k: -rates[k] for k in short})
    best_p99 = min(latencies[k][1] for k in timed)
    near = [k for k in timed if latencies[k][1] <= (1.0 + p99_tolerance) * best_p99]
    return pick_best({k: -rates[k] for k in near})


def tune_server(engine, n_requests=4096, candidates=None, latency_requests=4096, repeats=SERVER_REPEATS,
                log=print):
    """Server schedule: burst throughput (path-steps/s, best of `repeats`
    runs) of the first `n_requests` requests of the config-4 stream (FP32)
    under each candidate, then p50/p99 latency of a `latency_requests` stream
    with Poisson arrivals at 0.8x the best burst capacity under each candidate
    within SERVER_RATE_TOLERANCE of that capacity (pick_server chooses)."""
    from .engine import Server
    from . import workloads
    cs, ps, _ = workloads.config4_stream(n=n_requests, seed=2024, precision=_lib.FP32)
    total = workloads.path_steps(cs)
    saved, engine.tuning = engine.tuning, None
    try:
        with Server(engine) as srv:                                   # warm-up
            srv.replay(cs, ps, [0.0] * len(cs))
        rates = {}
        for cand in candidates or SERVER_CANDIDATES:
            key = tuple(sorted(cand.items()))
            for _ in range(repeats):                                  # best of the repeats
                with Server(engine, **cand) as srv:
                    t0 = time.perf_counter()
                    srv.replay(cs, ps, [0.0] * len(cs))
                    wall = time.perf_counter() - t0
                rates[key] = max(rates.get(key, 0.0), total / wall)
            log("server %s -> %.3e path-steps/s" % (cand, rates[key]))
        latencies = {}
        if latency_requests:
            top = max(rates.values())
            cl, pl, _ = workloads.config4_stream(n=latency_requests, seed=2025, precision=_lib.FP32)
            rate = 0.8 * top / (workloads.path_steps(cl) / len(cl))
            cl, pl, arr = workloads.config4_stream(n=latency_requests, seed=2025, precision=_lib.FP32,
                                                    rate=rate)
            for key, v in rates.items():
                if v < (1.0 - SERVER_RATE_TOLERANCE) * top:
                    continue
                with Server(engine, **dict(key)) as srv:
                    lat = sorted(r.latency_s for r in srv.replay(cl, pl, arr))
                q = lambda f: lat[min(len(lat) - 1, int(f * len(lat)))] * 1e3
                latencies[key] = (q(0.50), q(0.99))
                log("server %s at %.0f req/s -> p50 %.1f ms, p99 %.1f ms"
                    % (dict(key), rate, latencies[key][0], latencies[key][1]))
    finally:
        engine.tuning = saved
    best = pick_server(rates, latencies)
    return (dict(best), {json.dumps(dict(k)): v for k, v in rates.items()},
            {json.dumps(dict(k)): [round(x, 3) for x in v] for k, v in latencies.items()})


def tune_server_only(engine, path: str = DEFAULT_PATH, log=print) -> dict:
    """Re-tunes the server schedule of the table at `path`, keeping its lane
    entries."""
    table = load(path)
    if table is None:
        raise FileNotFoundError(path)
    server, rates, lat = tune_server(engine, log=log)
    table.update(server=server, server_path_steps_per_s=rates, server_latency_ms=lat,
                 created=time.strftime("%Y-%m-%d %H:%M:%S"))
    validate(table)
    with open(path, "w") as f:
        json.dump(table, f, indent=1)
    return table


def autotune(engine, out: str = DEFAULT_PATH, quick: bool = False, log=print) -> dict:
    if quick:
        lanes = tune_lanes(engine, precisions=(_lib.FP32,), schemes=(_lib.MILSTEIN,), lane_counts=(4, 13),
                           n_paths=1 << 16, n_steps=32, repeats=1, log=log)
        server, rates, lat = tune_server(engine, n_requests=64, candidates=SERVER_CANDIDATES[:2],
                                         latency_requests=64, repeats=1, log=log)
    else:
        lanes = tune_lanes(engine, log=log)
        server, rates, lat = tune_server(engine, log=log)
    table = {"arch": ARCH, "lanes": lanes, "server": server, "server_path_steps_per_s": rates,
             "server_latency_ms": lat, "created": time.strftime("%Y-%m-%d %H:%M:%S")}
    validate(table)
    if out:
        os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
        with open(out, "w") as f:
            json.dump(table, f, indent=1)
    return table


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=DEFAULT_PATH)
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--server-only", action="store_true",
                    help="re-tune the server schedule of the table at --out, keeping its lane entries")
    ap.add_argument("--device", type=int, default=0)
    args = ap.parse_args()
    from .engine import Engine
    with Engine(args.device, tuning=None) as eng:
        log = lambda m: print(m, flush=True)
        if args.server_only:
            tune_server_only(eng, args.out, log=log)
        else:
            autotune(eng, args.out, args.quick, log=log)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
