"""The oracle's Greek stencils (qko_greeks_eval, MCEngine.hpp:586-730) pinned in
magnitude by the Black-Scholes limit.

With kappa = theta = sigma = rho = 0 the variance stays at v0, so each bumped
repricing is an unbiased Monte Carlo price of a Black-Scholes contract at
volatility sqrt(v0), and the expectation of every finite-difference Greek is
the same stencil applied to Black-Scholes prices.  That differs from the
closed-form Greek only by the stencil's truncation error, at most 0.25 % of the
Greek at the bumps used here (vomma; delta, rho 3e-5, gamma, vega 2e-4,
theta 5e-4, vanna 9e-4).  Tolerance: 5 standard errors of the mean over 8
independent seeds plus 0.5 % of the closed form.  A stencil with the wrong
denominator ((2h)^2 for h^2), a bump of v0 instead of sqrt(v0), or a missing
price scale is off by a factor of 2 or more, far outside it.
"""
import math

import numpy as np
import pytest

from bs_closed_form import call_greeks, cash_binary_call_greeks

S0, T, R, Q, V0 = 100.0, 1.0, 0.03, 0.01, 0.04
SEEDS = range(1000, 1008)
CALL_KEYS = ("price", "delta", "gamma", "vega", "vomma", "vanna", "rho", "theta")
# The cash binary's gamma, vomma and vanna are second differences of an
# indicator: far too noisy to pin at this path count.
BINARY_KEYS = ("price", "delta", "vega", "rho", "theta")


def bs_model(oracle, scheme, seed, n_steps):
    return oracle.model(model=oracle.HESTON, scheme=scheme, n_steps=n_steps, seed=seed, S0=1.0, v0=V0, r=R,
                        q=Q, T=T, kappa=0.0, theta=0.0, sigma=0.0, rho=0.0)


def seed_runs(oracle, scheme, payoff, n_paths, n_steps, bumps):
    out = []
    for sd in SEEDS:
        # price scale S0 for every payoff: a cash binary's price is not scaled
        # (OptionPricer.hpp:323-332), its spot bump still is (MCEngine.hpp:590-595)
        g = oracle.greeks(bs_model(oracle, scheme, sd, n_steps), n_paths, payoff, bumps, S0)
        out.append({"price": g.base_price, **{k: getattr(g, k) for k in CALL_KEYS[1:]}})
    return out


def check(runs, ref, keys):
    for k in keys:
        a = np.array([r[k] for r in runs])
        se = a.std(ddof=1) / math.sqrt(len(a))
        tol = 5.0 * se + 0.005 * abs(ref[k])
        assert abs(a.mean() - ref[k]) <= tol, (k, a.mean(), ref[k], se)


@pytest.mark.parametrize("scheme", [0, 1])
def test_call_greeks_black_scholes_limit(oracle, scheme):
    K = 110.0   # out of the money: d1, d2 != 0, so vomma and vanna are not zero
    b = oracle.Bumps(1, 1, 1, 1, 1, 0, 0.01, 0.01, 1e-4, 1.0 / 365.0)
    runs = seed_runs(oracle, scheme, oracle.Payoff(oracle.EUROPEAN, oracle.CALL, K / S0, 1.0), 200_000, 4, b)
    check(runs, call_greeks(S0, K, T, R, Q, math.sqrt(V0)), CALL_KEYS)


def test_cash_binary_greeks_black_scholes_limit(oracle):
    K = 105.0
    b = oracle.Bumps(1, 1, 0, 1, 1, 0, 0.01, 0.01, 1e-4, 1.0 / 365.0)
    runs = seed_runs(oracle, 1, oracle.Payoff(oracle.BINARY_CASH, oracle.CALL, K / S0, 1.0), 200_000, 4, b)
    check(runs, cash_binary_call_greeks(S0, K, T, R, Q, math.sqrt(V0)), BINARY_KEYS)


def test_central_theta_black_scholes_limit(oracle):
    """theta_central adds the T + h repricing: (P(T - h) - P(T + h)) / 2h."""
    K = 110.0
    b = oracle.Bumps(0, 0, 0, 0, 1, 1, 0.01, 0.01, 1e-4, 1.0 / 365.0)
    runs = seed_runs(oracle, 0, oracle.Payoff(oracle.EUROPEAN, oracle.CALL, K / S0, 1.0), 200_000, 4, b)
    check(runs, call_greeks(S0, K, T, R, Q, math.sqrt(V0)), ("theta",))
