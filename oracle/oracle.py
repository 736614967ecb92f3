"""ctypes wrapper of the CPU oracle (oracle/qk_oracle.c).

Test infrastructure only: tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg import this module, as the checker.  The product package
(quantkos_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "libqk_oracle.so")

HESTON, BATES = 0, 1
EULER, MILSTEIN, QE = 0, 1, 2
(EUROPEAN, ASIAN, BARRIER_UO, BARRIER_DO, BARRIER_UI, BARRIER_DI, LOOKBACK_FLOAT,
 LOOKBACK_FIXED, BINARY_CASH, BINARY_ASSET, AMERICAN) = range(11)
CALL, PUT = 0, 1


class Model(C.Structure):
    _fields_ = [("model", C.c_int32), ("scheme", C.c_int32), ("fp32_semantics", C.c_int32),
                ("n_steps", C.c_int32), ("seed", C.c_uint64),
                ("S0", C.c_double), ("v0", C.c_double), ("r", C.c_double), ("q", C.c_double),
                ("T", C.c_double), ("kappa", C.c_double), ("theta", C.c_double),
                ("sigma", C.c_double), ("rho", C.c_double), ("lambda_j", C.c_double),
                ("mu_j", C.c_double), ("sigma_j", C.c_double)]


class Payoff(C.Structure):
    _fields_ = [("type", C.c_int32), ("right", C.c_int32), ("K", C.c_double), ("B", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("price", C.c_double), ("se", C.c_double), ("prob_itm", C.c_double),
                ("mean", C.c_double), ("m2", C.c_double), ("n", C.c_int64), ("itm", C.c_int64)]


class Bumps(C.Structure):
    _fields_ = [("delta_gamma", C.c_int32), ("vega_vomma", C.c_int32), ("vanna", C.c_int32),
                ("rho", C.c_int32), ("theta", C.c_int32), ("theta_central", C.c_int32),
                ("rel_S", C.c_double), ("abs_vol", C.c_double), ("abs_r", C.c_double),
                ("abs_T", C.c_double)]


class Greeks(C.Structure):
    _fields_ = [("delta", C.c_double), ("gamma", C.c_double), ("vega", C.c_double),
                ("vomma", C.c_double), ("vanna", C.c_double), ("rho", C.c_double),
                ("theta", C.c_double), ("base_price", C.c_double), ("base_se", C.c_double),
                ("n_pricings", C.c_int32)]


def build() -> str:
    """Compiles the oracle with its Makefile (gcc) and returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        P = C.POINTER
        d = C.c_double
        L.qko_philox4x32_10.argtypes = [P(C.c_uint32), P(C.c_uint32), P(C.c_uint32)]
        L.qko_uniform.argtypes = [C.c_uint32, C.c_int]
        L.qko_uniform.restype = d
        L.qko_normals.argtypes = [C.c_uint64, C.c_int, C.c_int64, C.c_int32, P(d)]
        L.qko_simulate.argtypes = [P(Model), C.c_int64, C.c_int64] + [P(d)] * 6
        L.qko_payoff_eval.argtypes = [P(Payoff), d, d, d, d, C.c_int32]
        L.qko_payoff_eval.restype = d
        L.qko_price.argtypes = [P(Model), C.c_int64, P(Payoff), C.c_int32, d, P(Stats)]
        L.qko_price_window.argtypes = [P(Model), C.c_int64, C.c_int64, P(Payoff), C.c_int32, d, P(Stats)]
        L.qko_price_american.argtypes = [P(Model), C.c_int64, C.c_int32, d, d, P(Stats), P(d), P(d)]
        L.qko_greeks_eval.argtypes = [P(Model), C.c_int64, P(Payoff), P(Bumps), d, P(Greeks)]
        L.qko_chan_merge.argtypes = [P(d), P(d), P(d), d, d, d]
        L.qko_lsm_solve.argtypes = [P(d), P(d)]
        L.qko_laguerre4.argtypes = [d, P(d)]
        L.qko_lsm_solve_n.argtypes = [P(d), C.c_int32, P(d)]
        L.qko_basis.argtypes = [d, C.c_int32, C.c_int32, P(d)]
        L.qko_price_american_ex.argtypes = [P(Model), C.c_int64, C.c_int32, d, C.c_int32, C.c_int32, d,
                                            P(Stats), P(d), P(d)]
        L.qko_gl64.argtypes = [P(d), P(d)]
        L.qko_heston_call.argtypes = [d] * 10
        L.qko_heston_call.restype = d
        L.qko_bates_call.argtypes = [d] * 13
        L.qko_bates_call.restype = d
        L.qko_bs_call.argtypes = [d] * 6
        L.qko_bs_call.restype = d
        _lib = L
    return _lib


def _dptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def model(**kw) -> Model:
    """Builds a Model; unspecified fields take the values of the reference's
    Config.hpp defaults where it has them."""
    m = Model()
    defaults = dict(model=HESTON, scheme=EULER, fp32_semantics=0, n_steps=5, seed=184467440737095,
                    S0=1.0, v0=0.04, r=0.05, q=0.0, T=1.0, kappa=2.0, theta=0.04, sigma=0.3,
                    rho=-0.7, lambda_j=0.1, mu_j=-0.1, sigma_j=0.15)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(m, k, v)
    return m


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().qko_philox4x32_10(c, k, o)
    return list(o)


def normals(seed, n_paths, n_steps, u24=False):
    z = np.empty((n_paths, n_steps, 2), dtype=np.float64)
    lib().qko_normals(seed, int(u24), n_paths, n_steps, _dptr(z))
    return z


def simulate(m: Model, n_paths: int, path0: int = 0, store=False):
    """Returns dict of per-path arrays: S_T, v_T, sum, max, min (and paths[step][path])."""
    out = {k: np.empty(n_paths) for k in ("S_T", "v_T", "sum", "max", "min")}
    paths = np.empty((m.n_steps, n_paths)) if store else None
    rc = lib().qko_simulate(C.byref(m), path0, n_paths, _dptr(paths), _dptr(out["S_T"]),
                            _dptr(out["v_T"]), _dptr(out["sum"]), _dptr(out["max"]),
                            _dptr(out["min"]))
    if rc != 0:
        raise ValueError("qko_simulate failed")
    if store:
        out["paths"] = paths
    return out


def payoff_eval(p: Payoff, S_T, S_sum, S_max, S_min, n_steps):
    return lib().qko_payoff_eval(C.byref(p), S_T, S_sum, S_max, S_min, n_steps)


def price(m: Model, n_paths: int, payoffs, price_scale: float = 1.0):
    arr = (Payoff * len(payoffs))(*payoffs)
    out = (Stats * len(payoffs))()
    rc = lib().qko_price(C.byref(m), n_paths, arr, len(payoffs), price_scale, out)
    if rc != 0:
        raise ValueError("qko_price failed")
    return list(out)


def price_window(m: Model, path0: int, n_paths: int, payoffs, price_scale: float = 1.0):
    """price() over the paths [path0, path0 + n_paths) of the stream."""
    arr = (Payoff * len(payoffs))(*payoffs)
    out = (Stats * len(payoffs))()
    rc = lib().qko_price_window(C.byref(m), path0, n_paths, arr, len(payoffs), price_scale, out)
    if rc != 0:
        raise ValueError("qko_price_window failed")
    return list(out)


def merge_stats(parts, disc: float, scale: float) -> Stats:
    """Chan's merge of per-window (n, mean, M2, itm) in window order, then the
    pricer's finalisation (OptionPricer.hpp:152-172, :204-249)."""
    n, mean, m2, itm = 0.0, 0.0, 0.0, 0
    for s in parts:
        n, mean, m2 = chan_merge((n, mean, m2), (float(s.n), s.mean, s.m2))
        itm += s.itm
    out = Stats()
    var = m2 / (n - 1.0) if n > 1 else 0.0
    out.price = mean * disc * scale
    out.se = (var / n) ** 0.5 * disc * scale
    out.prob_itm = itm / n
    out.mean, out.m2, out.n, out.itm = mean, m2, int(n), itm
    return out


def price_american(m: Model, n_paths: int, right: int, K: float, price_scale: float = 1.0,
                   want_coeffs=False, want_cf=False, n_basis=4, weighted=False):
    """Stats; with want_coeffs also the coefficients [n_steps][n_basis]; with
    want_cf also the final discounted cash flows per path (after the day-0
    rule).  n_basis 3/4/5 = LaguerreP02/P03/P04; weighted multiplies the basis
    by exp(-X/2)."""
    out = Stats()
    co = np.zeros((m.n_steps, n_basis)) if want_coeffs else None
    cf = np.zeros(n_paths) if want_cf else None
    rc = lib().qko_price_american_ex(C.byref(m), n_paths, right, K, n_basis, int(bool(weighted)), price_scale,
                                     C.byref(out), _dptr(co), _dptr(cf))
    if rc != 0:
        raise ValueError("qko_price_american failed")
    res = (out,) + ((co,) if want_coeffs else ()) + ((cf,) if want_cf else ())
    return res if len(res) > 1 else out


# Rank fractions of the payoff distribution (OptionPricer.hpp:345-357).
PERCENTILE_Q = (0.01, 0.05, 0.10, 0.20, 0.30, 0.40, 0.50, 0.60, 0.70, 0.80, 0.90, 0.95, 0.99)


def percentiles(values, factor=1.0):
    """analyze_risk_neutral_discounted_payoff_distribution (OptionPricer.hpp:340-360):
    sort ascending; p0 = first, p_q = element floor(n*q) (FP64 product,
    truncated), p100 = last; each times `factor` (discount x scale)."""
    v = np.sort(np.asarray(values, dtype=np.float64))
    n = len(v)
    ranks = [0] + [int(float(n) * q) for q in PERCENTILE_Q] + [n - 1]
    return [float(v[r]) * factor for r in ranks]


def greeks(m: Model, n_paths: int, p: Payoff, b: Bumps, price_scale: float = 1.0) -> Greeks:
    g = Greeks()
    rc = lib().qko_greeks_eval(C.byref(m), n_paths, C.byref(p), C.byref(b), price_scale,
                               C.byref(g))
    if rc != 0:
        raise ValueError("qko_greeks_eval failed")
    return g


def lsm_solve(sums):
    s = np.ascontiguousarray(sums, dtype=np.float64)
    c = np.zeros(4)
    fb = lib().qko_lsm_solve(_dptr(s), _dptr(c))
    return c, bool(fb)


def basis(X, n_basis=4, weighted=False):
    L = np.zeros(n_basis)
    lib().qko_basis(X, n_basis, int(bool(weighted)), _dptr(L))
    return L


def lsm_solve_n(sums, n_basis):
    """sums = [count, M (n_basis x n_basis, row-major), B (n_basis)] -> (C, fallback)."""
    s = np.ascontiguousarray(sums, dtype=np.float64)
    c = np.zeros(n_basis)
    fb = lib().qko_lsm_solve_n(_dptr(s), n_basis, _dptr(c))
    return c, bool(fb)


def laguerre4(X):
    L = np.zeros(4)
    lib().qko_laguerre4(X, _dptr(L))
    return L


def chan_merge(a, b):
    n, m, q = C.c_double(a[0]), C.c_double(a[1]), C.c_double(a[2])
    lib().qko_chan_merge(C.byref(n), C.byref(m), C.byref(q), *b)
    return n.value, m.value, q.value


def gl64():
    x, w = np.zeros(64), np.zeros(64)
    lib().qko_gl64(_dptr(x), _dptr(w))
    return x, w


def heston_call(S0, K, T, r, q, v0, kappa, theta, sigma, rho):
    return lib().qko_heston_call(S0, K, T, r, q, v0, kappa, theta, sigma, rho)


def bates_call(S0, K, T, r, q, v0, kappa, theta, sigma, rho, lam, mu_j, sigma_j):
    return lib().qko_bates_call(S0, K, T, r, q, v0, kappa, theta, sigma, rho, lam, mu_j, sigma_j)


def bs_call(S0, K, T, r, q, vol):
    return lib().qko_bs_call(S0, K, T, r, q, vol)
