"""Payoff distribution percentiles selected on the device (qk_price_distribution)
against the oracle's sort of the same per-path payoffs (OptionPricer.hpp:340-360).

Tolerances: FP64 relative 1e-9 (per-path payoffs differ from the oracle's by
libm vs device-math ulps; order statistics move by no more); FP32 Euler 2e-5
relative or 2e-4 absolute (FP32 path arithmetic: the spots carry ~1e-7 relative
rounding per step, and the payoff avg - K of a path near the money turns that
into an absolute error of price_scale x S x 1e-6 at most over 30 steps);
American FP64: the device's Gram sums are added in another order than the
oracle's, so an exercise decision within rounding of the continuation value
can go the other way on a few paths; each such path moves an order statistic
by at most one rank, so the American percentiles must lie within 3 ranks of
the oracle's sorted cash flows."""
import math

import numpy as np
import pytest

import quantkos_amd as qk
from quantkos_amd import engine as qe

pytestmark = pytest.mark.gpu

FIELDS = ("p0", "p1", "p5", "p10", "p20", "p30", "p40", "median", "p60", "p70", "p80", "p90", "p95", "p99", "p100")


def contract(precision, scheme, n_paths, n_steps=30, **kw):
    d = dict(model=qk.HESTON, scheme=scheme, precision=precision, n_steps=n_steps, n_paths=n_paths, seed=77,
             S0=1.0, v0=0.04, r=0.03, q=0.0, T=1.0, kappa=2.0, theta=0.04, sigma=0.5, rho=-0.7, price_scale=100.0)
    d.update(kw)
    return qe.contract(**d)


def omodel(o, c):
    return o.model(model=c.model, scheme=c.scheme, fp32_semantics=int(c.precision == qk.FP32),
                   n_steps=c.n_steps, seed=c.seed, S0=c.S0, v0=c.v0, r=c.r, q=c.q, T=c.T,
                   kappa=c.kappa, theta=c.theta, sigma=c.sigma, rho=c.rho, lambda_j=c.lambda_j,
                   mu_j=c.mu_j, sigma_j=c.sigma_j)


def oracle_percentiles(o, c, p):
    sim = o.simulate(omodel(o, c), c.n_paths)
    op = o.Payoff(p.type, p.right, p.K, p.B)
    x = np.array([o.payoff_eval(op, sim["S_T"][i], sim["sum"][i], sim["max"][i], sim["min"][i], c.n_steps)
                  for i in range(c.n_paths)])
    factor = math.exp(-c.r * c.T) * (1.0 if p.type == qk.BINARY_CASH else c.price_scale)
    return o.percentiles(x, factor)


PAYOFFS = [qe.payoff(qk.EUROPEAN, qk.CALL, 1.0), qe.payoff(qk.ASIAN, qk.PUT, 1.0),
           qe.payoff(qk.BARRIER_UP_OUT, qk.CALL, 1.0, 1.25), qe.payoff(qk.LOOKBACK_FLOATING, qk.PUT, 1.0),
           qe.payoff(qk.BINARY_CASH, qk.CALL, 1.05)]


@pytest.mark.parametrize("p", PAYOFFS, ids=lambda p: "type%d" % p.type)
def test_fp64_percentiles_match_oracle(engine, oracle, p):
    c = contract(qk.FP64, qk.MILSTEIN, 20_011)
    s, d = engine.price_distribution(c, p)
    ref = oracle_percentiles(oracle, c, p)
    got = [getattr(d, f) for f in FIELDS]
    for f, a, b in zip(FIELDS, got, ref):
        assert math.isclose(a, b, rel_tol=1e-9, abs_tol=1e-12), (f, a, b)
    plain = engine.price(c, p)
    assert s.price == plain.price and s.itm == plain.itm
    assert all(x <= y for x, y in zip(got, got[1:]))


def test_fp32_percentiles_match_oracle(engine, oracle):
    p = qe.payoff(qk.ASIAN, qk.CALL, 1.0)
    c = contract(qk.FP32, qk.EULER, 8_192)
    _, d = engine.price_distribution(c, p)
    ref = oracle_percentiles(oracle, c, p)
    for f, b in zip(FIELDS, ref):
        assert math.isclose(getattr(d, f), b, rel_tol=2e-5, abs_tol=2e-4), (f, getattr(d, f), b)


def test_american_percentiles_match_oracle(engine, oracle):
    c = contract(qk.FP64, qk.MILSTEIN, 20_000, n_steps=25, T=0.5, sigma=0.3, rho=-0.5, r=0.05)
    p = qe.payoff(qk.AMERICAN, qk.PUT, 1.05)
    s, d = engine.price_distribution(c, p)
    o, cf = oracle.price_american(omodel(oracle, c), c.n_paths, qk.PUT, 1.05, 100.0, want_cf=True)
    v = np.sort(cf) * c.price_scale
    n = len(v)
    ranks = [0] + [int(float(n) * q) for q in oracle.PERCENTILE_Q] + [n - 1]
    for f, r in zip(FIELDS, ranks):
        lo, hi = v[max(r - 3, 0)], v[min(r + 3, n - 1)]
        assert lo * (1 - 1e-12) - 1e-12 <= getattr(d, f) <= hi * (1 + 1e-12) + 1e-12, (f, getattr(d, f), v[r])
    assert math.isclose(s.price, engine.price_american(c, qk.PUT, 1.05).price, rel_tol=1e-12)


def test_american_immediate_exercise_gives_a_point_mass(engine):
    c = contract(qk.FP64, qk.MILSTEIN, 5_000, n_steps=10, v0=1e-4, theta=1e-4, sigma=0.01, rho=0.0, r=0.10,
                 price_scale=1.0)
    s, d = engine.price_distribution(c, qe.payoff(qk.AMERICAN, qk.PUT, 2.0))
    assert s.price == 1.0 and s.se == 0.0
    assert all(getattr(d, f) == 1.0 for f in FIELDS)


@pytest.mark.parametrize("precision", [qk.FP32, qk.FP64])
def test_large_distribution_properties(engine, precision):
    """At the benchmark's size: ordered, bounded by the payoff's range, the median
    of a call below its mean (right skew), binary payoffs a two-point law."""
    c = contract(precision, qk.QE, 4_000_000, n_steps=52)
    _, d = engine.price_distribution(c, qe.payoff(qk.EUROPEAN, qk.CALL, 1.0))
    got = [getattr(d, f) for f in FIELDS]
    assert all(x <= y for x, y in zip(got, got[1:])) and got[0] == 0.0
    s, b = engine.price_distribution(c, qe.payoff(qk.BINARY_CASH, qk.CALL, 1.0))
    disc = math.exp(-c.r * c.T)
    vals = {round(getattr(b, f) / disc, 12) for f in FIELDS}
    assert vals <= {0.0, 1.0}
    # P(S_T > K) is s.prob_itm: the ranks below n(1 - P) are zero, those above one
    zero_frac = 1.0 - s.prob_itm
    for f, q in zip(FIELDS[1:-1], (0.01, 0.05, 0.10, 0.20, 0.30, 0.40, 0.50, 0.60, 0.70, 0.80, 0.90, 0.95, 0.99)):
        if abs(q - zero_frac) > 1e-3:
            assert getattr(b, f) == (0.0 if q < zero_frac else disc), (f, q, zero_frac)
