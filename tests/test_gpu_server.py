"""The continuous-batching server on the GPU.

Tolerances: a request's result is bit-identical alone and in any mix (tile
moments are stored per request and merged in tile order); against the plain
single-contract launch it agrees to 1e-12 relative in FP64 and 1e-5 in FP32
(the same arithmetic in a different kernel; FMA contraction may differ).
"""
import copy
import math

import pytest

import quantkos_amd as qk
from quantkos_amd import workloads
from quantkos_amd.engine import QKError, Server

pytestmark = pytest.mark.gpu


def small_stream(precision, n=24, seed=7):
    cs, ps, _ = workloads.config4_stream(n=n, seed=seed, precision=precision, paths_range=(2e3, 4e4))
    return cs, ps


def plain(engine, c, p):
    if p.type == qk.AMERICAN:
        return engine.price_american(c, p.right, p.K)
    return engine.price(c, p)


def same(a, b):
    return (a.price, a.se, a.itm, a.n) == (b.price, b.se, b.itm, b.n)


@pytest.mark.parametrize("precision", [qk.FP64, qk.FP32])
def test_server_matches_plain_launch(engine, precision):
    cs, ps = small_stream(precision)
    with Server(engine) as srv:
        res = srv.replay(cs, ps, [0.0] * len(cs))
    tol = 1e-12 if precision == qk.FP64 else 1e-5
    for c, p, r in zip(cs, ps, res):
        assert r.status == 0
        ref = plain(engine, c, p)
        assert r.stats.n == c.n_paths
        assert math.isclose(r.stats.price, ref.price, rel_tol=tol, abs_tol=1e-9), (p.type, r.stats.price, ref.price)
        assert math.isclose(r.stats.se, ref.se, rel_tol=max(tol, 1e-9), abs_tol=1e-9)
        assert r.latency_s >= r.queue_s >= 0.0


@pytest.mark.parametrize("precision", [qk.FP64, qk.FP32])
def test_server_step_remainders_match_plain_launch(engine, precision):
    """The batch kernel's step loops for every remainder of the Philox block
    sharing (QE: four steps per three blocks; Euler and Milstein: two steps per
    block), for path-dependent and European payoffs, against the plain launch."""
    cs, ps = [], []
    for i, n in enumerate((1, 2, 3, 5, 6, 7, 50, 101)):
        for scheme in (qk.QE, qk.EULER, qk.MILSTEIN):
            for model in (qk.HESTON, qk.BATES):
                cs.append(qk.contract(model=model, scheme=scheme, precision=precision, n_steps=n,
                                      n_paths=3_000 + 64 * i + 7, seed=100 + i, lambda_j=0.5))
                ps.append(qk.payoff(qk.ASIAN if i % 2 else qk.EUROPEAN, qk.CALL, 1.0))
    with Server(engine) as srv:
        res = srv.replay(cs, ps, [0.0] * len(cs))
    tol = 1e-12 if precision == qk.FP64 else 1e-5
    for c, p, r in zip(cs, ps, res):
        assert r.status == 0
        ref = plain(engine, c, p)
        assert r.stats.n == c.n_paths, (c.n_steps, c.scheme, c.model)
        if precision == qk.FP64:
            assert r.stats.itm == ref.itm, (c.n_steps, c.scheme, c.model)
        assert math.isclose(r.stats.price, ref.price, rel_tol=tol, abs_tol=1e-9), (c.n_steps, c.scheme, c.model)


def test_same_result_alone_and_in_mix(engine):
    cs, ps = small_stream(qk.FP32, n=20, seed=11)
    with Server(engine) as srv:
        alone = []
        for c, p in zip(cs, ps):
            alone.append(srv.wait(srv.submit(c, p)).stats)
    # the same requests interleaved with others, arriving together and staggered
    extra_c, extra_p = small_stream(qk.FP32, n=30, seed=99)
    mixed_c = cs + extra_c
    mixed_p = ps + extra_p
    order = list(range(len(mixed_c)))
    order = order[::3] + order[1::3] + order[2::3]
    with Server(engine, max_inflight=2, item_path_steps=20_000) as srv:
        res = srv.replay([mixed_c[i] for i in order], [mixed_p[i] for i in order],
                         [0.0005 * k for k in range(len(order))])
    pos = {i: k for k, i in enumerate(order)}
    for i in range(len(cs)):
        r = res[pos[i]]
        assert r.status == 0 and same(r.stats, alone[i]), (i, r.stats.price, alone[i].price)


def test_memory_admission_queues_and_rejects(engine):
    c = qk.contract(model=qk.HESTON, scheme=qk.QE, precision=qk.FP32, n_steps=100, n_paths=40_000, seed=5)
    am = qk.payoff(qk.AMERICAN, qk.PUT, 1.05)
    need = qk._lib.load().qk_footprint_bytes(c, 1, 1, 1)
    # an arena that holds one American request at a time: the rest wait for memory
    with Server(engine, mem_budget_bytes=int(need * 1.5)) as srv:
        tickets = [srv.submit(copy.copy(c), am) for _ in range(4)]
        out = [srv.wait(t) for t in tickets]
        st = srv.stats()
        assert all(r.status == 0 for r in out)
        assert len({r.stats.price for r in out}) == 1
        assert st.arena_peak_bytes <= st.arena_bytes
        big = copy.copy(c)
        big.n_paths = 400_000
        r = srv.wait(srv.submit(big, am))
        assert r.status == qk._lib.QK_ENOMEM


def test_ticket_semantics(engine):
    c = qk.contract(n_paths=10_000, n_steps=50, seed=3)
    with Server(engine) as srv:
        t = srv.submit(c, qk.payoff(qk.EUROPEAN, qk.CALL, 1.0))
        r = srv.wait(t)
        assert r.status == 0
        with pytest.raises(QKError):
            srv.wait(t)
        with pytest.raises(QKError):
            srv.submit(qk.contract(n_paths=0), qk.payoff(qk.EUROPEAN))
        t2 = srv.submit(c, qk.payoff(qk.ASIAN, qk.PUT, 1.0))
        got = None
        for _ in range(100000):
            got = srv.poll(t2)
            if got is not None:
                break
        assert got is not None and got.status == 0


def test_config4_stream_poisson(engine):
    """A reduced config-4 stream (paths 1e4..2e5) with Poisson arrivals."""
    cs, ps, arr = workloads.config4_stream(n=200, seed=3, paths_range=(1e4, 2e5), rate=2000.0)
    with Server(engine) as srv:
        res = srv.replay(cs, ps, arr)
        st = srv.stats()
    assert all(r.status == 0 for r in res)
    assert st.n_completed == 200 and st.path_steps == workloads.path_steps(cs)
    assert all(math.isfinite(r.stats.price) and r.stats.price >= 0 for r in res)
    lat = sorted(r.latency_s for r in res)
    assert lat[len(lat) // 2] > 0


def oracle_model(o, c):
    return o.model(model=c.model, scheme=c.scheme, fp32_semantics=int(c.precision == qk.FP32),
                   n_steps=c.n_steps, seed=c.seed, S0=c.S0, v0=c.v0, r=c.r, q=c.q, T=c.T, kappa=c.kappa,
                   theta=c.theta, sigma=c.sigma, rho=c.rho, lambda_j=c.lambda_j, mu_j=c.mu_j,
                   sigma_j=c.sigma_j)


def test_server_mix_matches_oracle(engine, oracle):
    """A mixed FP64 batch (every payoff family, Heston and Bates QE, American by
    Longstaff-Schwartz) against the FP64 oracle on the same random numbers: the
    reference prices each request by one MCDispatcher run (Distpatcher.hpp:78-83,
    ForwardMC / ForwardBackwardMC.hpp:204-255).  Non-American prices within 1e-9
    relative with exact ITM counts; American prices within 0.02 standard errors
    (the early-date Laguerre regressions are badly conditioned, so summation
    order moves single coefficients, not the fitted policy)."""
    cs, ps = small_stream(qk.FP64, n=40, seed=21)
    with Server(engine) as srv:
        res = srv.replay(cs, ps, [0.0005 * k for k in range(len(cs))])
    kinds = set()
    for c, p, r in zip(cs, ps, res):
        assert r.status == 0
        m = oracle_model(oracle, c)
        kinds.add(p.type)
        if p.type == qk.AMERICAN:
            ref = oracle.price_american(m, c.n_paths, p.right, p.K, c.price_scale)
            assert abs(r.stats.price - ref.price) <= 0.02 * ref.se + 1e-12, (r.stats.price, ref.price, ref.se)
        else:
            ref = oracle.price(m, c.n_paths, [oracle.Payoff(p.type, p.right, p.K, p.B)], c.price_scale)[0]
            assert math.isclose(r.stats.price, ref.price, rel_tol=1e-9, abs_tol=1e-12), (p.type, r.stats.price,
                                                                                          ref.price)
            assert r.stats.itm == ref.itm and r.stats.n == ref.n
            assert math.isclose(r.stats.se, ref.se, rel_tol=1e-7, abs_tol=1e-12)
    assert qk.AMERICAN in kinds and qk.EUROPEAN in kinds and len(kinds) >= 6, kinds


def test_config4_stream_full_size(engine):
    """The config-4 stream at its size: 4,096 requests, paths log-uniform in
    1e4..1e7, steps in {12, 52, 252, 504}, the whole mix, burst-submitted to the
    server in FP32 and again in FP64 (same random numbers).  Every request
    completes, the server counts exactly the stream's path-steps, and every
    non-American FP32 price is within 1 standard error of its FP64 price."""
    cs32, ps, _ = workloads.config4_stream(n=4096, seed=2024, precision=qk.FP32)
    cs64, _, _ = workloads.config4_stream(n=4096, seed=2024, precision=qk.FP64)
    out = {}
    for name, cs in (("fp32", cs32), ("fp64", cs64)):
        with Server(engine) as srv:
            res = srv.replay(cs, ps, [0.0] * len(cs))
            st = srv.stats()
        assert st.n_completed == len(cs) and st.path_steps == workloads.path_steps(cs), name
        assert all(r.status == 0 for r in res), name
        assert all(r.stats.n == c.n_paths for r, c in zip(res, cs)), name
        out[name] = res
    worst = 0.0
    for p, a, b in zip(ps, out["fp32"], out["fp64"]):
        if p.type == qk.AMERICAN or b.stats.se == 0.0:
            continue
        worst = max(worst, abs(a.stats.price - b.stats.price) / b.stats.se)
    assert worst <= 1.0, worst


@pytest.mark.parametrize("precision", [qk.FP64, qk.FP32])
def test_server_american_bases_match_plain_launch(engine, precision):
    """American requests with each regression basis (qk_server_submit_ex) in one
    mix: the pump advances each (precision, basis) group with its own date
    kernel, and every result agrees with qk_price_american_ex's for that basis.
    The server's pump splits the paths into other blocks than the plain call,
    so the Gram sums differ in rounding; with the worse-conditioned bases that
    can flip an exercise decision near the continuation value, hence 0.05
    standard errors there and the plain-launch tolerance for LaguerreP03."""
    cases = [(qk.LAGUERRE_P02, False), (qk.LAGUERRE_P03, False), (qk.LAGUERRE_P04, False),
             (qk.LAGUERRE_P03, True), (qk.LAGUERRE_P04, True)]
    # a contract whose degree-4 Gram stays well conditioned at every date
    # (v0 = theta = 0.25, K = 1.5 S0, T = 2; see the other-bases oracle test)
    cs = [qk.contract(model=qk.BATES, scheme=qk.QE, precision=precision, n_steps=30 + 7 * i, n_paths=20_000 + 999 * i,
                      seed=50 + i, S0=1.0, v0=0.25, r=0.05, T=2.0, kappa=2.0, theta=0.25, sigma=0.3, rho=-0.7)
          for i in range(len(cases))]
    p = qk.payoff(qk.AMERICAN, qk.PUT, 1.5)
    with Server(engine) as srv:
        tickets = [srv.submit(c, p, basis=b, weighted=w) for c, (b, w) in zip(cs, cases)]
        res = [srv.wait(t) for t in tickets]
    tol = 1e-12 if precision == qk.FP64 else 1e-5
    for c, (b, w), r in zip(cs, cases, res):
        assert r.status == 0
        ref = engine.price_american(c, qk.PUT, 1.5, basis=b, weighted=w)
        if b == qk.LAGUERRE_P03 and not w:
            assert math.isclose(r.stats.price, ref.price, rel_tol=tol, abs_tol=1e-12), (r.stats.price, ref.price)
        else:
            assert abs(r.stats.price - ref.price) <= 0.05 * ref.se, (b, w, r.stats.price, ref.price, ref.se)


@pytest.mark.parametrize("precision", [qk.FP32, qk.FP64])
def test_batch_cap_and_american_first_do_not_change_results(engine, precision):
    """A batch work cap of one item per resident wavefront (a batch per few
    requests, American requests admitted ahead of earlier arrivals) and a
    small arena (American requests wait for memory) give the same bits as
    the default schedule: a request's result depends only on (seed, path).
    The same holds with a second batch admitted only while the queue holds
    two batches' work (overlap_min_backlog), without the express lane, and
    with an express lane that takes nearly every non-American request."""
    cs, ps = small_stream(precision, n=40, seed=23)
    with Server(engine) as srv:
        ref = srv.replay(cs, ps, [0.0] * len(cs))
    rs = 4 if precision == qk.FP32 else 8
    need = max(c.n_paths * c.n_steps * rs for c, p in zip(cs, ps) if p.type == qk.AMERICAN)
    for kw in (dict(overlap_min_backlog=0, express_max_path_steps=-1),
               dict(overlap_min_backlog=2, express_max_path_steps=-1),
               dict(overlap_min_backlog=0, express_max_path_steps=1 << 34)):
        with Server(engine, batch_items_per_wave=1, blocks_per_cu=2, max_inflight=2, item_path_steps=4096,
                    mem_budget_bytes=2 * need + (64 << 20), **kw) as srv:
            res = srv.replay(cs, ps, [0.0] * len(cs))
            st = srv.stats()
        # one work item per wavefront makes many batches; an express lane that
        # takes every non-American request may need only one batch beside them
        assert st.n_batches > (2 if kw["express_max_path_steps"] < 0 else 1), (kw, st.n_batches)
        assert st.n_completed == len(cs), kw
        for p, a, b in zip(ps, ref, res):
            assert a.status == 0 and b.status == 0
            assert same(a.stats, b.stats), (kw, p.type, a.stats.price, b.stats.price)


@pytest.mark.parametrize("precision", [qk.FP64, qk.FP32])
def test_every_payoff_type_through_the_server(engine, precision):
    """Each of the ten payoff types, call and put, submitted together to the
    server (normal batches and the express lane) equals its plain launch (FP64
    to 1e-12 with equal ITM counts, FP32 to 1e-5)."""
    types = (qk.EUROPEAN, qk.ASIAN, qk.BARRIER_UP_OUT, qk.BARRIER_DOWN_OUT, qk.BARRIER_UP_IN,
             qk.BARRIER_DOWN_IN, qk.LOOKBACK_FLOATING, qk.LOOKBACK_FIXED, qk.BINARY_CASH, qk.BINARY_ASSET)
    cs, ps = [], []
    for i, t in enumerate(types):
        for right in (qk.CALL, qk.PUT):
            B = 1.2 if t in (qk.BARRIER_UP_OUT, qk.BARRIER_UP_IN) else 0.85
            cs.append(qk.contract(model=qk.HESTON if i % 2 else qk.BATES, scheme=qk.QE, precision=precision,
                                  n_steps=52, n_paths=20_000 + 448 * i + 64 * right + 5, seed=300 + 2 * i + right,
                                  lambda_j=0.3))
            ps.append(qk.payoff(t, right, 1.05 if right == qk.CALL else 0.95, B))
    with Server(engine) as srv:
        res = srv.replay(cs, ps, [0.0] * len(cs))
    tol = 1e-12 if precision == qk.FP64 else 1e-5
    for c, p, r in zip(cs, ps, res):
        assert r.status == 0, (p.type, p.right)
        ref = plain(engine, c, p)
        assert r.stats.n == c.n_paths, (p.type, p.right)
        if precision == qk.FP64:
            assert r.stats.itm == ref.itm, (p.type, p.right)
        assert math.isclose(r.stats.price, ref.price, rel_tol=tol, abs_tol=1e-12), (p.type, p.right, r.stats.price,
                                                                                     ref.price)
