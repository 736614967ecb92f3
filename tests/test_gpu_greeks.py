"""GPU Greeks (qk_price_greeks: the reference's stencils MCEngine.hpp:586-730 as
lanes of one launch) pinned in magnitude, FP32 against FP64, and the FP32
lanes-per-thread kernels against single launches.

Black-Scholes limit: with kappa = theta = sigma = rho = 0 the variance stays at
v0 (Euler, Milstein: v' = v), so every lane prices a Black-Scholes contract at
volatility sqrt(v0) without bias, and the expectation of each finite-difference
Greek is the stencil applied to Black-Scholes prices: the closed-form Greek up
to the stencil's truncation, below 0.25 % of it at these bumps.  Tolerance: 5
standard errors of the mean over 8 independent seeds plus 0.5 % of the closed
form.  A wrong stencil denominator, a bump of v0 instead of sqrt(v0), or a
missing price scale misses by a factor of 2 or more.
"""
import contextlib
import math

import numpy as np
import pytest

import quantkos_amd as qk
from quantkos_amd import autotune
from quantkos_amd import engine as qe
from bs_closed_form import call_greeks, cash_binary_call_greeks

pytestmark = pytest.mark.gpu

S0, T, R, Q, V0 = 100.0, 1.0, 0.03, 0.01, 0.04
CALL_KEYS = ("base_price", "delta", "gamma", "vega", "vomma", "vanna", "rho", "theta")
BINARY_KEYS = ("base_price", "delta", "vega", "rho", "theta")
# Config-3 bumps: S0 +-1 %, sqrt(v0) +-1e-3, r +-1bp, T -+1/365 (central): 13 lanes.
CONFIG3_BUMPS = dict(rel_S=0.01, abs_vol=1e-3, abs_r=1e-4, abs_T=1.0 / 365.0, theta_central=1)


@contextlib.contextmanager
def schedule(engine, entries):
    """Installs a lanes-per-thread table for the block, then restores the
    engine's own (its tuning table, or none)."""
    engine.set_schedule(entries)
    try:
        yield
    finally:
        if engine.tuning:
            autotune.apply(engine, engine.tuning)
        else:
            engine.set_schedule([])


def bs_contract(precision, seed, n_paths, n_steps=252):
    return qe.contract(model=qk.HESTON, scheme=qk.MILSTEIN, precision=precision, n_steps=n_steps,
                       n_paths=n_paths, seed=seed, S0=1.0, v0=V0, r=R, q=Q, T=T, kappa=0.0, theta=0.0,
                       sigma=0.0, rho=0.0, price_scale=S0)


@pytest.mark.parametrize("precision", [qk.FP64, qk.FP32])
def test_greeks_black_scholes_limit(engine, precision):
    pays = [qe.payoff(qk.EUROPEAN, qk.CALL, 1.10), qe.payoff(qk.BINARY_CASH, qk.CALL, 1.05)]
    b = qe.bumps(**CONFIG3_BUMPS)
    # FP32 runs the 13-lane stencil at 8 lanes per thread, as the gfx950 table does
    with schedule(engine, [(qk.FP32, qk.MILSTEIN, 13, 8)]):
        assert engine.lanes_per_thread(qk.FP32, qk.MILSTEIN, 13) == 8
        runs = [engine.greeks(bs_contract(precision, 1000 + k, 500_000), pays, b) for k in range(8)]
    assert runs[0][0].n_lanes == 13
    vol = math.sqrt(V0)
    for j, (ref, keys) in enumerate(((call_greeks(S0, 110.0, T, R, Q, vol), CALL_KEYS),
                                     (cash_binary_call_greeks(S0, 105.0, T, R, Q, vol), BINARY_KEYS))):
        for k in keys:
            a = np.array([getattr(r[j], k) for r in runs])
            want = ref["price" if k == "base_price" else k]
            se = a.std(ddof=1) / math.sqrt(len(a))
            assert abs(a.mean() - want) <= 5.0 * se + 0.005 * abs(want), (j, k, a.mean(), want, se)


@pytest.mark.parametrize("scheme", [qk.EULER, qk.MILSTEIN, qk.QE])
def test_fp32_lanes_per_thread_match_single_launches(engine, scheme):
    """13 bumped FP32 lanes at 1, 2, 4 and 8 lanes per thread (8 leaves a
    remainder group of 4 and 1) against one launch per lane; 61 steps leave a
    remainder of the Philox block sharing."""
    base = qe.contract(model=qk.HESTON, scheme=scheme, precision=qk.FP32, n_steps=61, n_paths=30_000,
                       seed=7, S0=1.0, v0=0.04, r=0.03, q=0.01, T=1.0, kappa=2.0, theta=0.04, sigma=0.5,
                       rho=-0.7)
    lanes = []
    for i in range(13):
        c = qe.contract(**{f: getattr(base, f) for f, _ in base._fields_})
        c.S0 = 1.0 + 0.01 * (i % 3 - 1)
        c.v0 = 0.04 + 0.002 * (i % 5)
        c.r = 0.03 + 0.001 * (i % 2)
        c.T = 1.0 - 0.01 * (i % 4)
        lanes.append(c)
    pays = [qe.payoff(qk.EUROPEAN, qk.CALL, 1.0), qe.payoff(qk.BARRIER_UP_OUT, qk.CALL, 1.0, 1.25),
            qe.payoff(qk.BINARY_CASH, qk.PUT, 0.95)]
    single = [engine.price_payoffs(c, pays) for c in lanes]
    n = base.n_paths
    for lpt in (1, 2, 4, 8):
        with schedule(engine, [(qk.FP32, scheme, 13, lpt)]):
            assert engine.lanes_per_thread(qk.FP32, scheme, 13) == lpt
            multi = engine.price_lanes(lanes, pays)
        for i, (row, ref) in enumerate(zip(multi, single)):
            for j, (a, b) in enumerate(zip(row, ref)):
                if pays[j].type in (qk.BINARY_CASH, qk.BARRIER_UP_OUT):
                    # FP32 paths of the two kernels may differ in the last bit
                    # (another FMA contraction), which can move a path across
                    # the strike or the barrier: each such path moves the
                    # price by at most the payoff's jump there over n
                    jump = 1.0 if pays[j].type == qk.BINARY_CASH else pays[j].B - pays[j].K
                    assert abs(a.price - b.price) <= 2.0 * jump / n + 1e-9 * abs(b.price), (lpt, i, j, a.price,
                                                                                          b.price)
                    assert abs(a.itm - b.itm) <= 2, (lpt, i, j, a.itm, b.itm)
                else:
                    assert math.isclose(a.price, b.price, rel_tol=2e-6, abs_tol=1e-9), (lpt, i, j, a.price,
                                                                                       b.price)
                    assert abs(a.itm - b.itm) <= 1, (lpt, i, j, a.itm, b.itm)


def test_fp32_greeks_within_one_se_of_fp64(engine):
    """Config-3 shape: Heston implicit Milstein, 2,000,000 paths x 252 steps,
    European call K=100 and cash-or-nothing call K=105 from one path set, the
    13-lane stencil; FP32 paths (FP64 differencing) against the FP64 run on the
    same random numbers.  Each Greek's standard error is that of 8 independent
    FP64 sub-runs of 250,000 paths divided by sqrt(8); the FP32 Greek must be
    within 1 of them (north-star rule for prices, applied to every Greek)."""
    kw = dict(model=qk.HESTON, scheme=qk.MILSTEIN, n_steps=252, S0=1.0, v0=0.04, r=0.03, q=0.0, T=1.0,
              kappa=2.0, theta=0.04, sigma=0.5, rho=-0.7, price_scale=100.0)
    pays = [qe.payoff(qk.EUROPEAN, qk.CALL, 1.0), qe.payoff(qk.BINARY_CASH, qk.CALL, 1.05)]
    b = qe.bumps(**CONFIG3_BUMPS)
    g32 = engine.greeks(qe.contract(precision=qk.FP32, n_paths=2_000_000, seed=42, **kw), pays, b)
    g64 = engine.greeks(qe.contract(precision=qk.FP64, n_paths=2_000_000, seed=42, **kw), pays, b)
    subs = [engine.greeks(qe.contract(precision=qk.FP64, n_paths=250_000, seed=500 + k, **kw), pays, b)
            for k in range(8)]
    keys = ("base_price", "delta", "gamma", "vega", "vomma", "vanna", "rho", "theta")
    for j in range(2):
        for k in keys:
            se = np.std([getattr(s[j], k) for s in subs], ddof=1) / math.sqrt(8)
            d = getattr(g32[j], k) - getattr(g64[j], k)
            assert abs(d) <= se, (j, k, getattr(g32[j], k), getattr(g64[j], k), se)


def test_vol_bump_equal_to_vol_is_accepted(engine, oracle):
    """abs_vol = sqrt(v0) is valid in the reference (Config.hpp:518): the
    down-bump prices at v0 = 0.  The C ABI accepts it and matches the oracle's
    stencil on the same random numbers."""
    c = qe.contract(model=qk.HESTON, scheme=qk.MILSTEIN, precision=qk.FP64, n_steps=20, n_paths=20_000, seed=3,
                    S0=1.0, v0=0.04, r=0.03, q=0.0, T=1.0, kappa=2.0, theta=0.04, sigma=0.3, rho=-0.7,
                    price_scale=100.0)
    b = qe.bumps(delta_gamma=0, vanna=0, rho=0, theta=0, vega_vomma=1, abs_vol=0.2)
    g = engine.greeks(c, [qe.payoff(qk.EUROPEAN, qk.CALL, 1.0)], b)[0]
    m = oracle.model(model=0, scheme=1, n_steps=20, seed=3, S0=1.0, v0=0.04, r=0.03, q=0.0, T=1.0, kappa=2.0,
                     theta=0.04, sigma=0.3, rho=-0.7)
    og = oracle.greeks(m, 20_000, oracle.Payoff(0, 0, 1.0, 1.0), oracle.Bumps(0, 1, 0, 0, 0, 0, 0.01, 0.2, 1e-4,
                                                                              1 / 365), 100.0)
    assert math.isclose(g.vega, og.vega, rel_tol=1e-9) and math.isclose(g.vomma, og.vomma, rel_tol=1e-7)
