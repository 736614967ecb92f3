"""Python handle on the C ABI, used by the tests and bench.py.

Every call goes to libquantkos.so on the GPU; nothing here computes prices.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import (Bumps, Contract, Distribution, Greeks, Payoff, Result, ScheduleEntry, ServerConfig, ServerStats,
                   Stats, Timing)


class QKError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__("%s (%d): %s" % (_lib.load().qk_strerror(status).decode(), status, message))
        self.status = status


def contract(**kw) -> Contract:
    """A qk_contract.  Defaults: the reference's Config.hpp defaults where it has
    them (r=0.05, q=0, kappa=2, theta=0.04, sigma=0.3, rho=-0.7, Bates jumps
    0.1/-0.1/0.15, seed 184467440737095), S0 = 1 (normalised), FP32, Euler."""
    c = Contract()
    d = dict(model=_lib.HESTON, scheme=_lib.EULER, precision=_lib.FP32, n_steps=252, n_paths=100_000,
             seed=184467440737095, S0=1.0, v0=0.04, r=0.05, q=0.0, T=1.0, kappa=2.0, theta=0.04,
             sigma=0.3, rho=-0.7, lambda_j=0.1, mu_j=-0.1, sigma_j=0.15, price_scale=1.0)
    d.update(kw)
    for k, v in d.items():
        setattr(c, k, v)
    return c


def payoff(type_, right=_lib.CALL, K=1.0, B=1.0) -> Payoff:
    return Payoff(type_, right, K, B)


def bumps(**kw) -> Bumps:
    """qk_bumps with the reference's default bump sizes (Config.hpp:443-446)."""
    b = Bumps()
    d = dict(delta_gamma=1, vega_vomma=1, vanna=1, rho=1, theta=1, theta_central=0,
             rel_S=0.001, abs_vol=0.01, abs_r=0.0001, abs_T=1.0 / 365.0)
    d.update(kw)
    for k, v in d.items():
        setattr(b, k, v)
    return b


def analytic_european(c: Contract, right: int, K: float, quadrature=None) -> float:
    """Semi-analytic European price of the contract's model (host only; Heston /
    Bates characteristic function, Black-Scholes / Merton when sigma = 0), in
    the contract's price scale.  quadrature: None (panels of 50 until the
    integrand decays) or (phi_max, panel), e.g. (100, 100) for the
    reference's Heston rule."""
    L = _lib.load()
    out = C.c_double()
    q = None if quadrature is None else C.byref(_lib.CfQuadrature(*quadrature))
    st = L.qk_analytic_european(C.byref(c), right, K, q, C.byref(out))
    if st != 0:
        raise QKError(st, "qk_analytic_european")
    return out.value


def load_config(path: str):
    """Config::Parser::parse + apply_price_scaling of a QuantKos YAML file (no device work)."""
    L = _lib.load()
    spec = _lib.RunSpec()
    err = C.create_string_buffer(1024)
    st = L.qk_config_load(str(path).encode(), C.byref(spec), err, 1024)
    if st != 0:
        raise QKError(st, err.value.decode())
    return spec


def run_config(path: str, device: int = 0, quiet: bool = True):
    """MCDispatcher(Parser::parse(path)).launch_montecarlo() on `device`."""
    L = _lib.load()
    res = _lib.RunResult()
    err = C.create_string_buffer(2048)
    st = L.qk_run_config(str(path).encode(), device, int(quiet), C.byref(res), err, 2048)
    if st != 0:
        raise QKError(st, err.value.decode())
    return res


class Engine:
    """One qk_ctx: a device, a memory budget and a stream.

    `tuning`: a table written by quantkos_amd.autotune (a dict or a JSON file
    path); "default" loads the package's table for this device architecture
    when there is one, None applies nothing.  The table sets the lane kernels'
    lanes per thread and the request server's default schedule."""

    def __init__(self, device: int = 0, mem_budget_bytes: int = 0, tuning="default"):
        self.L = _lib.load()
        h = C.c_void_p()
        st = self.L.qk_ctx_create(device, mem_budget_bytes, C.byref(h))
        if st != 0:
            raise QKError(st, "qk_ctx_create(device=%d)" % device)
        self.h = h
        self.tuning = None
        if tuning is not None:
            from . import autotune
            table = autotune.load() if tuning == "default" else (
                tuning if isinstance(tuning, dict) else autotune.load(tuning))
            if table is not None:
                autotune.apply(self, table)

    def set_schedule(self, entries):
        """entries: iterable of (precision, scheme, n_lanes, lanes_per_thread)."""
        entries = list(entries)
        arr = (ScheduleEntry * max(1, len(entries)))(*[ScheduleEntry(*e) for e in entries])
        self._check(self.L.qk_set_schedule(self.h, arr, len(entries)))

    def set_lsm_schedule(self, mode: int):
        """LSM_AUTO, LSM_PER_DATE or LSM_COOPERATIVE for the backward pass."""
        self._check(self.L.qk_set_lsm_schedule(self.h, mode))

    def lanes_per_thread(self, precision: int, scheme: int, n_lanes: int) -> int:
        return self.L.qk_lanes_per_thread(self.h, precision, scheme, n_lanes)

    def close(self):
        if self.h:
            self.L.qk_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st):
        if st != 0:
            raise QKError(st, self.L.qk_last_error(self.h).decode())

    def price(self, c: Contract, p: Payoff) -> Stats:
        out = Stats()
        self._check(self.L.qk_price(self.h, C.byref(c), C.byref(p), C.byref(out)))
        return out

    def price_payoffs(self, c: Contract, payoffs) -> list:
        arr = (Payoff * len(payoffs))(*payoffs)
        out = (Stats * len(payoffs))()
        self._check(self.L.qk_price_payoffs(self.h, C.byref(c), arr, len(payoffs), out))
        return list(out)

    def price_payoffs_noalloc(self, c: Contract, payoffs) -> list:
        """price_payoffs through the allocation-free entry point (fixed
        per-device scratch; QK_ENOMEM beyond it)."""
        arr = (Payoff * len(payoffs))(*payoffs)
        out = (Stats * len(payoffs))()
        self._check(self.L.qk_price_payoffs_noalloc(self.h, C.byref(c), arr, len(payoffs), out))
        return list(out)

    def price_window(self, c: Contract, path0: int, payoffs) -> list:
        """Paths [path0, path0 + c.n_paths) of the path set: per payoff, the
        window's merge chunks (one Moments per 65,536 paths)."""
        nc = -(-c.n_paths // _lib.WINDOW_CHUNK_PATHS)
        arr = (Payoff * len(payoffs))(*payoffs)
        out = (_lib.Moments * (len(payoffs) * nc))()
        self._check(self.L.qk_price_window(self.h, C.byref(c), path0, arr, len(payoffs), out))
        return [list(out[j * nc:(j + 1) * nc]) for j in range(len(payoffs))]

    def merge_windows(self, c: Contract, payoffs, chunks) -> list:
        """Merges per-payoff chunk lists (windows concatenated in path order)
        into one Stats per payoff, discounted and scaled as c prices."""
        n = len(chunks[0])
        assert all(len(x) == n for x in chunks) and len(chunks) == len(payoffs)
        arr = (Payoff * len(payoffs))(*payoffs)
        flat = (_lib.Moments * (len(payoffs) * n))(*[m for x in chunks for m in x])
        out = (Stats * len(payoffs))()
        self._check(self.L.qk_merge_windows(self.h, C.byref(c), arr, len(payoffs), flat, n, out))
        return list(out)

    def price_lanes(self, lanes, payoffs) -> list:
        la = (Contract * len(lanes))(*lanes)
        pa = (Payoff * len(payoffs))(*payoffs)
        out = (Stats * (len(lanes) * len(payoffs)))()
        self._check(self.L.qk_price_lanes(self.h, la, len(lanes), pa, len(payoffs), out))
        n = len(payoffs)
        return [list(out[i * n:(i + 1) * n]) for i in range(len(lanes))]

    def greeks(self, c: Contract, payoffs, b: Bumps) -> list:
        pa = (Payoff * len(payoffs))(*payoffs)
        out = (Greeks * len(payoffs))()
        self._check(self.L.qk_price_greeks(self.h, C.byref(c), pa, len(payoffs), C.byref(b), out))
        return list(out)

    def price_american(self, c: Contract, right: int, K: float, want_coeffs: bool = False,
                       basis: int = _lib.LAGUERRE_P03, weighted: bool = False):
        """Longstaff-Schwartz price; basis LAGUERRE_P02/P03/P04 (3/4/5 functions),
        weighted multiplies them by exp(-X/2).  want_coeffs adds [n_steps][n_basis]."""
        out = Stats()
        nb = 3 + basis
        co = np.zeros((c.n_steps, nb)) if want_coeffs else None
        ptr = co.ctypes.data_as(C.POINTER(C.c_double)) if want_coeffs else None
        if basis == _lib.LAGUERRE_P03 and not weighted:
            self._check(self.L.qk_price_american(self.h, C.byref(c), right, K, C.byref(out), ptr))
        else:
            opt = _lib.LsmOptions(basis, int(bool(weighted)))
            self._check(self.L.qk_price_american_ex(self.h, C.byref(c), right, K, C.byref(opt), C.byref(out), ptr))
        return (out, co) if want_coeffs else out

    def price_distribution(self, c: Contract, p: Payoff):
        """(Stats, Distribution): the price and the discounted payoff percentiles
        of the same paths; p.type AMERICAN prices by Longstaff-Schwartz."""
        out, d = Stats(), Distribution()
        self._check(self.L.qk_price_distribution(self.h, C.byref(c), C.byref(p), C.byref(out), C.byref(d)))
        return out, d

    def simulate_paths(self, c: Contract, path0: int = 0) -> np.ndarray:
        """Spot paths [path0, path0 + c.n_paths) laid out [step][path]."""
        S = np.empty((c.n_steps, c.n_paths), dtype=np.float64)
        self._check(self.L.qk_simulate_paths(self.h, C.byref(c), path0,
                                             S.ctypes.data_as(C.POINTER(C.c_double))))
        return S

    def price_store_paths(self, c: Contract, p: Payoff) -> Stats:
        """qk_price that keeps its path set on the device (for stored_paths)."""
        out = Stats()
        self._check(self.L.qk_price_store_paths(self.h, C.byref(c), C.byref(p), C.byref(out)))
        return out

    def stored_paths(self, path0: int, n: int, n_steps: int, precision: int) -> np.ndarray:
        """Paths [path0, path0 + n) of the last stored path set, [path][step],
        float32 for FP32 runs and float64 for FP64 runs."""
        a = np.empty((n, n_steps), dtype=np.float32 if precision == _lib.FP32 else np.float64)
        self._check(self.L.qk_copy_stored_paths(self.h, path0, n, a.ctypes.data_as(C.c_void_p)))
        return a

    def timing(self) -> Timing:
        t = Timing()
        self._check(self.L.qk_last_timing(self.h, C.byref(t)))
        return t

    def synchronize(self):
        self._check(self.L.qk_synchronize(self.h))


class Server:
    """A continuous-batching request server (qk_server_*) on an Engine's device.

    The keyword arguments are the fields of qk_server_config (include/qk.h);
    a 0 (or None for the last two) takes the engine's tuning table entry, then
    the library default.  Every schedule gives the same bits for a request:
    results depend only on the request itself."""

    def __init__(self, engine: Engine, mem_budget_bytes: int = 0, max_inflight: int = 0,
                 max_lsm_streams: int = 0, item_path_steps: int = 0, blocks_per_cu: int = 0,
                 batch_items_per_wave: int = 0, overlap_min_backlog: int = None,
                 express_max_path_steps: int = None):
        self.engine = engine
        self.L = engine.L
        tuned = (engine.tuning or {}).get("server") or {}
        max_inflight = max_inflight or int(tuned.get("max_inflight", 0))
        item_path_steps = item_path_steps or int(tuned.get("item_path_steps", 0))
        blocks_per_cu = blocks_per_cu or int(tuned.get("blocks_per_cu", 0))
        batch_items_per_wave = batch_items_per_wave or int(tuned.get("batch_items_per_wave", 0))
        if overlap_min_backlog is None:       # 0 is a setting of its own (overlap whenever a stream is free)
            overlap_min_backlog = int(tuned.get("overlap_min_backlog", 0))
        if express_max_path_steps is None:    # 0: the library's default lane; < 0: none
            express_max_path_steps = int(tuned.get("express_max_path_steps", 0))
        cfg = ServerConfig(mem_budget_bytes, max_inflight, max_lsm_streams, item_path_steps, blocks_per_cu,
                           batch_items_per_wave, overlap_min_backlog, express_max_path_steps)
        h = C.c_void_p()
        engine._check(self.L.qk_server_create(engine.h, C.byref(cfg), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.L.qk_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, c: Contract, p: Payoff, basis: int = _lib.LAGUERRE_P03, weighted: bool = False) -> int:
        """Ticket of a request; basis / weighted choose the regression basis of
        an American request (LaguerreP03 unweighted by default)."""
        t = C.c_uint64()
        if basis == _lib.LAGUERRE_P03 and not weighted:
            self.engine._check(self.L.qk_server_submit(self.h, C.byref(c), C.byref(p), C.byref(t)))
        else:
            opt = _lib.LsmOptions(basis, int(bool(weighted)))
            self.engine._check(self.L.qk_server_submit_ex(self.h, C.byref(c), C.byref(p), C.byref(opt),
                                                          C.byref(t)))
        return t.value

    def poll(self, ticket: int):
        """Result, or None while pending."""
        r = Result()
        st = self.L.qk_server_poll(self.h, ticket, C.byref(r))
        if st == _lib.QK_ENOTREADY:
            return None
        if st != 0:
            raise QKError(st, "qk_server_poll(%d)" % ticket)
        return r

    def wait(self, ticket: int) -> Result:
        r = Result()
        st = self.L.qk_server_wait(self.h, ticket, C.byref(r))
        if st != 0:
            raise QKError(st, "qk_server_wait(%d)" % ticket)
        return r

    def replay(self, contracts, payoffs, arrival_s) -> list:
        n = len(contracts)
        ca = (Contract * n)(*contracts)
        pa = (Payoff * n)(*payoffs)
        ta = (C.c_double * n)(*arrival_s)
        out = (Result * n)()
        st = self.L.qk_server_replay(self.h, ca, pa, ta, n, out)
        if st != 0:
            raise QKError(st, "qk_server_replay")
        return list(out)

    def stats(self) -> ServerStats:
        s = ServerStats()
        self.engine._check(self.L.qk_server_stats_get(self.h, C.byref(s)))
        return s
