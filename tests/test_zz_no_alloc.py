"""qk_price_payoffs_noalloc (the allocation-free entry point) against
qk_price_payoffs through the C ABI: the same kernels over the same tile
partition, so the results are bit-identical.  Two calls in a row, the second
smaller, so stale contents of the shared array would show; a call beyond the
fixed array fails with QK_ENOMEM."""
import ctypes as C

import pytest

import quantkos_amd as qk
from quantkos_amd import _lib
from quantkos_amd._lib import Payoff, Stats

pytestmark = pytest.mark.gpu


def fields(s):
    return (s.price, s.se, s.prob_itm, s.mean, s.m2, s.n, s.itm)


def call_noalloc(engine, c, pays):
    arr = (Payoff * len(pays))(*pays)
    out = (Stats * len(pays))()
    st = engine.L.qk_price_payoffs_noalloc(engine.h, C.byref(c), arr, len(pays), out)
    return st, list(out)


@pytest.mark.parametrize("precision", [qk.FP32, qk.FP64])
def test_noalloc_matches_allocating_entry(engine, precision):
    pays = [qk.payoff(qk.ASIAN, qk.CALL, 1.0), qk.payoff(qk.BARRIER_UP_OUT, qk.CALL, 1.0, 1.3),
            qk.payoff(qk.LOOKBACK_FLOATING, qk.PUT, 1.0)]
    big = qk.contract(model=qk.HESTON, scheme=qk.QE, precision=precision, n_steps=100, n_paths=1_000_003,
                      seed=42, S0=1.0, v0=0.04, r=0.03, T=1.0, kappa=2.0, theta=0.04, sigma=0.5, rho=-0.7,
                      price_scale=100.0)
    small = qk.contract(model=qk.BATES, scheme=qk.EULER, precision=precision, n_steps=37, n_paths=70_001,
                        seed=9, S0=1.0, v0=0.05, r=0.02, T=0.5, price_scale=100.0)
    for c, ps in ((big, pays), (small, pays[:2])):
        st, got = call_noalloc(engine, c, ps)
        assert st == 0, engine.L.qk_last_error(engine.h)
        ref = engine.price_payoffs(c, ps)
        for a, b in zip(got, ref):
            assert fields(a) == fields(b), (fields(a), fields(b))


def test_noalloc_rejects_beyond_its_array(engine):
    pays = [qk.payoff(qk.EUROPEAN, qk.CALL, 1.0)] * 4
    # 4 payoffs x 20 B per 64 paths: 16 MiB covers about 13M paths
    c = qk.contract(scheme=qk.EULER, precision=qk.FP32, n_steps=4, n_paths=20_000_000, seed=1)
    st, _ = call_noalloc(engine, c, pays)
    assert st == _lib.QK_ENOMEM
    assert b"scratch" in engine.L.qk_last_error(engine.h)
    st, got = call_noalloc(engine, qk.contract(scheme=qk.EULER, n_steps=4, n_paths=5_000, seed=1), pays[:1])
    assert st == 0 and got[0].n == 5_000
