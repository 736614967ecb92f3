"""Batches as path windows (qk_price_window + qk_merge_windows): the reference
prices a path set batch by batch (run_forward_all_mc_batches, MCEngine.hpp:
489-520) and accumulates every batch into one estimator (OptionPricer.hpp:
123-193).  Windows that start at multiples of 65,536 paths reproduce the
one-launch merge exactly; other 64-aligned windows merge the same moments in
another order.  Each window is also checked against the oracle's own window."""
import copy
import math

import pytest

import quantkos_amd as qk
from quantkos_amd import engine as qe

pytestmark = pytest.mark.gpu

PAYS = [qe.payoff(qk.ASIAN, qk.CALL, 1.0), qe.payoff(qk.BARRIER_UP_OUT, qk.CALL, 1.0, 1.3),
        qe.payoff(qk.LOOKBACK_FLOATING, qk.PUT, 1.0)]


def contract(precision, n_paths, **kw):
    d = dict(model=qk.HESTON, scheme=qk.QE, precision=precision, n_steps=24, n_paths=n_paths, seed=42,
             S0=1.0, v0=0.04, r=0.03, q=0.0, T=1.0, kappa=2.0, theta=0.04, sigma=0.5, rho=-0.7,
             price_scale=100.0)
    d.update(kw)
    return qe.contract(**d)


def windowed(engine, c, windows, pays=PAYS):
    chunks = [[] for _ in pays]
    for p0, n in windows:
        cw = copy.copy(c)
        cw.n_paths = n
        for j, x in enumerate(engine.price_window(cw, p0, pays)):
            assert len(x) == -(-n // qk.WINDOW_CHUNK_PATHS)
            chunks[j] += x
    return engine.merge_windows(c, pays, chunks)


def key(s):
    return (s.price, s.se, s.prob_itm, s.mean, s.m2, s.n, s.itm)


@pytest.mark.parametrize("precision", [qk.FP64, qk.FP32])
def test_chunk_aligned_windows_equal_one_launch(engine, precision):
    """300,000 paths in windows of 131,072 + 131,072 + 37,856 (a ragged last
    window, a partial last tile) against qk_price_payoffs: every field equal."""
    c = contract(precision, 300_000)
    ref = engine.price_payoffs(c, PAYS)
    got = windowed(engine, c, [(0, 131_072), (131_072, 131_072), (262_144, 37_856)])
    for a, b in zip(got, ref):
        assert key(a) == key(b), (key(a), key(b))


def test_one_window_equals_one_launch(engine):
    """A single window of one chunk: the second merge stage copies it unchanged."""
    c = contract(qk.FP64, 50_000, model=qk.BATES, lambda_j=0.5, mu_j=-0.05, sigma_j=0.1)
    ref = engine.price_payoffs(c, PAYS)
    got = windowed(engine, c, [(0, 50_000)])
    for a, b in zip(got, ref):
        assert key(a) == key(b)


def test_unaligned_windows_merge_exactly(engine):
    """Windows at multiples of 64 but not of 65,536 hold the same paths: counts
    equal, moments equal to rounding."""
    c = contract(qk.FP64, 200_000)
    ref = engine.price_payoffs(c, PAYS)
    got = windowed(engine, c, [(0, 70_016), (70_016, 64), (70_080, 129_920)])
    for a, b in zip(got, ref):
        assert (a.n, a.itm) == (b.n, b.itm)
        assert math.isclose(a.price, b.price, rel_tol=1e-12)
        assert math.isclose(a.se, b.se, rel_tol=1e-9)


@pytest.mark.parametrize("path0", [0, 65_536 * 3, 2 ** 32 - 64])
def test_window_matches_oracle_window(engine, oracle, path0):
    """A window of 5,000 paths priced alone equals the oracle's same window of
    the stream (FP64 1e-9 relative, ITM counts exact), also across the 2^32
    path-index boundary."""
    c = contract(qk.FP64, 5_000)
    chunks = engine.price_window(c, path0, PAYS)
    got = engine.merge_windows(c, PAYS, chunks)
    m = oracle.model(model=c.model, scheme=c.scheme, n_steps=c.n_steps, seed=c.seed, S0=c.S0, v0=c.v0, r=c.r,
                     q=c.q, T=c.T, kappa=c.kappa, theta=c.theta, sigma=c.sigma, rho=c.rho)
    o = oracle.price_window(m, path0, c.n_paths, [oracle.Payoff(p.type, p.right, p.K, p.B) for p in PAYS], 100.0)
    for g, x in zip(got, o):
        assert math.isclose(g.price, x.price, rel_tol=1e-9, abs_tol=1e-12), (g.price, x.price)
        assert g.itm == x.itm


def test_window_contract_errors(engine):
    """path0 off a 64-path tile, negative path0 and an empty chunk list are
    QK_EINVAL; the ctx stays usable."""
    import ctypes as C
    c = contract(qk.FP32, 1_000)
    for p0 in (37, -64):
        with pytest.raises(qe.QKError):
            engine.price_window(c, p0, PAYS)
    arr = (qk._lib.Payoff * 1)(PAYS[0])
    flat = (qk._lib.Moments * 1)()
    out = (qk._lib.Stats * 1)()
    assert engine.L.qk_merge_windows(engine.h, C.byref(c), arr, 1, flat, 0, out) == qk._lib.QK_EINVAL
    assert engine.price(c, PAYS[0]).n == 1_000
