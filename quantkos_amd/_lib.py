"""ctypes binding of libquantkos.so (include/qk.h).

The library is the product: there is no Python or CPU fallback.  Loading fails
loudly if the in-tree build is missing.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libquantkos.so")

# The request server runs up to two batch streams, an express-lane stream and
# four Longstaff-Schwartz pump streams beside the ctx stream.  HIP maps
# streams onto a per-process pool
# of hardware queues (4 by default); streams beyond the pool share a queue, and
# a pump stream that shares a batch stream's queue waits behind every batch
# kernel (measured on the 4,096-request stream: that chain's dates ~10x slower,
# American p99 latency 224-344 ms instead of ~75 ms).  Twelve queues give every
# stream its own with room to spare, so a smaller setting (or none) is raised
# to 12 unless QK_KEEP_HW_QUEUES=1.  HIP reads the setting when it
# initialises, so it takes effect when this package is imported before
# anything else in the process uses the GPU.
MIN_HW_QUEUES = 12
try:
    _hwq = int(os.environ.get("GPU_MAX_HW_QUEUES", "0"))
except ValueError:
    _hwq = 0
if _hwq < MIN_HW_QUEUES and os.environ.get("QK_KEEP_HW_QUEUES") != "1":
    os.environ["GPU_MAX_HW_QUEUES"] = str(MIN_HW_QUEUES)

QK_OK, QK_EINVAL, QK_ENOMEM, QK_EDEVICE, QK_EAGAIN, QK_ENOTFOUND, QK_ENOTREADY = 0, -1, -2, -3, -4, -5, -6
HESTON, BATES = 0, 1
EULER, MILSTEIN, QE = 0, 1, 2
(EUROPEAN, ASIAN, BARRIER_UP_OUT, BARRIER_DOWN_OUT, BARRIER_UP_IN, BARRIER_DOWN_IN,
 LOOKBACK_FLOATING, LOOKBACK_FIXED, BINARY_CASH, BINARY_ASSET, AMERICAN) = range(11)
CALL, PUT = 0, 1
FP32, FP64 = 0, 1
LAGUERRE_P02, LAGUERRE_P03, LAGUERRE_P04 = 0, 1, 2
LSM_AUTO, LSM_PER_DATE, LSM_COOPERATIVE = 0, 1, 2
NOALLOC_SCRATCH_BYTES = 16 << 20   # QK_NOALLOC_SCRATCH_BYTES
WINDOW_CHUNK_PATHS = 65536         # QK_WINDOW_CHUNK_PATHS


class Contract(C.Structure):
    _fields_ = [("model", C.c_int32), ("scheme", C.c_int32), ("precision", C.c_int32),
                ("n_steps", C.c_int32), ("n_paths", C.c_int64), ("seed", C.c_uint64),
                ("S0", C.c_double), ("v0", C.c_double), ("r", C.c_double), ("q", C.c_double),
                ("T", C.c_double), ("kappa", C.c_double), ("theta", C.c_double),
                ("sigma", C.c_double), ("rho", C.c_double), ("lambda_j", C.c_double),
                ("mu_j", C.c_double), ("sigma_j", C.c_double), ("price_scale", C.c_double)]


class Payoff(C.Structure):
    _fields_ = [("type", C.c_int32), ("right", C.c_int32), ("K", C.c_double), ("B", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("price", C.c_double), ("se", C.c_double), ("prob_itm", C.c_double),
                ("mean", C.c_double), ("m2", C.c_double), ("n", C.c_int64), ("itm", C.c_int64)]


class Moments(C.Structure):
    _fields_ = [("n", C.c_double), ("mean", C.c_double), ("m2", C.c_double), ("itm", C.c_uint64)]


class Bumps(C.Structure):
    _fields_ = [("delta_gamma", C.c_int32), ("vega_vomma", C.c_int32), ("vanna", C.c_int32),
                ("rho", C.c_int32), ("theta", C.c_int32), ("theta_central", C.c_int32),
                ("rel_S", C.c_double), ("abs_vol", C.c_double), ("abs_r", C.c_double),
                ("abs_T", C.c_double)]


class Greeks(C.Structure):
    _fields_ = [("delta", C.c_double), ("gamma", C.c_double), ("vega", C.c_double),
                ("vomma", C.c_double), ("vanna", C.c_double), ("rho", C.c_double),
                ("theta", C.c_double), ("base_price", C.c_double), ("base_se", C.c_double),
                ("n_lanes", C.c_int32)]


class Timing(C.Structure):
    _fields_ = [("path_kernel_ms", C.c_double), ("reduce_ms", C.c_double),
                ("lsm_backward_ms", C.c_double), ("total_ms", C.c_double),
                ("path_steps", C.c_int64)]


class Distribution(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("p0", "p1", "p5", "p10", "p20", "p30", "p40", "median", "p60", "p70",
                                          "p80", "p90", "p95", "p99", "p100", "select_s")]


class LsmOptions(C.Structure):
    _fields_ = [("basis", C.c_int32), ("weighted", C.c_int32)]


class RunSpec(C.Structure):
    _fields_ = [("contract", Contract), ("payoff", Payoff), ("bumps", Bumps),
                ("compute_greeks", C.c_int32), ("write_paths", C.c_int32), ("lsm", LsmOptions),
                ("analyze_distribution", C.c_int32)]


class RunResult(C.Structure):
    _fields_ = [("stats", Stats), ("greeks", Greeks), ("t_total_s", C.c_double),
                ("t_pricing_s", C.c_double), ("t_greeks_s", C.c_double), ("dist", Distribution),
                ("analytic_price", C.c_double), ("analytic_err_se", C.c_double)]


class CfQuadrature(C.Structure):
    _fields_ = [("phi_max", C.c_double), ("panel", C.c_double)]


class ServerConfig(C.Structure):
    _fields_ = [("mem_budget_bytes", C.c_size_t), ("max_inflight", C.c_int32),
                ("max_lsm_streams", C.c_int32), ("item_path_steps", C.c_int64),
                ("blocks_per_cu", C.c_int32), ("batch_items_per_wave", C.c_int32),
                ("overlap_min_backlog", C.c_int32), ("express_max_path_steps", C.c_int64)]


class ScheduleEntry(C.Structure):
    _fields_ = [("precision", C.c_int32), ("scheme", C.c_int32), ("n_lanes", C.c_int32),
                ("lanes_per_thread", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("stats", Stats), ("latency_s", C.c_double), ("queue_s", C.c_double),
                ("status", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("n_submitted", C.c_int64), ("n_completed", C.c_int64), ("n_batches", C.c_int64),
                ("path_steps", C.c_int64), ("arena_bytes", C.c_int64),
                ("arena_peak_bytes", C.c_int64), ("busy_s", C.c_double)]


_lib = None


def load():
    """Loads the in-tree libquantkos.so and declares its C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libquantkos.so is not built (%s); run `python -m quantkos_amd.build`"
                           % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    vp = C.c_void_p
    st = C.c_int
    sig = {
        "qk_ctx_create": ([C.c_int, C.c_size_t, P(vp)], st),
        "qk_ctx_destroy": ([vp], st),
        "qk_strerror": ([C.c_int], C.c_char_p),
        "qk_last_error": ([vp], C.c_char_p),
        "qk_synchronize": ([vp], st),
        "qk_last_timing": ([vp, P(Timing)], st),
        "qk_footprint_bytes": ([P(Contract), C.c_int32, C.c_int32, C.c_int32], C.c_size_t),
        "qk_set_schedule": ([vp, P(ScheduleEntry), C.c_int32], st),
        "qk_set_lsm_schedule": ([vp, C.c_int32], st),
        "qk_lanes_per_thread": ([vp, C.c_int32, C.c_int32, C.c_int32], C.c_int32),
        "qk_price": ([vp, P(Contract), P(Payoff), P(Stats)], st),
        "qk_price_payoffs": ([vp, P(Contract), P(Payoff), C.c_int32, P(Stats)], st),
        "qk_price_payoffs_noalloc": ([vp, P(Contract), P(Payoff), C.c_int32, P(Stats)], st),
        "qk_price_window": ([vp, P(Contract), C.c_int64, P(Payoff), C.c_int32, P(Moments)], st),
        "qk_merge_windows": ([vp, P(Contract), P(Payoff), C.c_int32, P(Moments), C.c_int64, P(Stats)], st),
        "qk_price_lanes": ([vp, P(Contract), C.c_int32, P(Payoff), C.c_int32, P(Stats)], st),
        "qk_price_greeks": ([vp, P(Contract), P(Payoff), C.c_int32, P(Bumps), P(Greeks)], st),
        "qk_price_american": ([vp, P(Contract), C.c_int32, C.c_double, P(Stats), P(C.c_double)], st),
        "qk_simulate_paths": ([vp, P(Contract), C.c_int64, P(C.c_double)], st),
        "qk_price_store_paths": ([vp, P(Contract), P(Payoff), P(Stats)], st),
        "qk_copy_stored_paths": ([vp, C.c_int64, C.c_int64, vp], st),
        "qk_price_distribution": ([vp, P(Contract), P(Payoff), P(Stats), P(Distribution)], st),
        "qk_price_american_ex": ([vp, P(Contract), C.c_int32, C.c_double, P(LsmOptions), P(Stats),
                                  P(C.c_double)], st),
        "qk_server_create": ([vp, P(ServerConfig), P(vp)], st),
        "qk_server_destroy": ([vp], st),
        "qk_server_submit": ([vp, P(Contract), P(Payoff), P(C.c_uint64)], st),
        "qk_server_submit_ex": ([vp, P(Contract), P(Payoff), P(LsmOptions), P(C.c_uint64)], st),
        "qk_server_poll": ([vp, C.c_uint64, P(Result)], st),
        "qk_server_wait": ([vp, C.c_uint64, P(Result)], st),
        "qk_server_stats_get": ([vp, P(ServerStats)], st),
        "qk_server_replay": ([vp, P(Contract), P(Payoff), P(C.c_double), C.c_int32, P(Result)], st),
        "qk_analytic_european": ([P(Contract), C.c_int32, C.c_double, P(CfQuadrature), P(C.c_double)], st),
        "qk_config_load": ([C.c_char_p, P(RunSpec), C.c_char_p, C.c_size_t], st),
        "qk_run_config": ([C.c_char_p, C.c_int, C.c_int32, P(RunResult), C.c_char_p, C.c_size_t], st),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L
