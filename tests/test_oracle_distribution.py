"""The oracle's payoff-distribution rule (OptionPricer.hpp:340-360) and the
cash flows its Longstaff-Schwartz restatement exposes for it."""
import numpy as np

import oracle.oracle as O


def test_percentile_ranks_on_a_known_grid():
    v = np.random.default_rng(3).permutation(1000).astype(float)
    assert O.percentiles(v) == [0, 10, 50, 100, 200, 300, 400, 500, 600, 700, 800, 900, 950, 990, 999]


def test_percentile_ranks_truncate_the_fp64_product():
    # n = 7: floor(7 q) = 0,0,0,1,2,2,3,4,4,5,6,6,6 for q = 0.01 ... 0.99 (7 * 0.7 = 4.8999999999999995)
    v = np.arange(7.0)[::-1]
    assert O.percentiles(v, 2.0) == [2.0 * x for x in (0, 0, 0, 0, 1, 2, 2, 3, 4, 4, 5, 6, 6, 6, 6)]


def test_percentiles_of_a_constant_and_a_two_point_law():
    assert O.percentiles(np.full(11, 3.5)) == [3.5] * 15
    v = np.r_[np.zeros(600), np.ones(400)]            # binary payoffs: 60 % zeros
    p = O.percentiles(v)
    assert p[:8] == [0.0] * 8 and p[8] == 1.0 and p[-1] == 1.0   # ranks 0..500 zero, 600 -> one


def test_american_cash_flows_match_the_price():
    m = O.model(model=0, scheme=1, n_steps=25, seed=5, S0=1.0, v0=0.04, r=0.05, q=0.0, T=0.5,
                kappa=2.0, theta=0.04, sigma=0.3, rho=-0.5)
    s, cf = O.price_american(m, 4000, 1, 1.05, 100.0, want_cf=True)
    assert cf.shape == (4000,) and (cf >= 0).all()
    assert abs(cf.mean() * 100.0 - s.price) < 1e-9 * s.price
    assert int((cf > 1e-12).sum()) == s.itm


def test_american_day0_rule_sets_every_cash_flow():
    # deep in the money, tiny volatility: exercising today beats holding
    m = O.model(model=0, scheme=1, n_steps=10, seed=5, S0=1.0, v0=1e-4, r=0.10, q=0.0, T=1.0,
                kappa=1.0, theta=1e-4, sigma=0.01, rho=0.0)
    s, cf = O.price_american(m, 500, 1, 2.0, 1.0, want_cf=True)
    assert np.all(cf == 1.0) and s.price == 1.0 and s.se == 0.0
