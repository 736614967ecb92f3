"""Pins the oracle's path algorithm, payoff semantics, statistics and LSM
against the reference's known answers and independent computations."""
import math

import numpy as np
import pytest

MODELS = [(0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (1, 2)]


def M(oracle, **kw):
    d = dict(model=0, scheme=0, n_steps=50, seed=42, S0=1.0, v0=0.04, r=0.03, q=0.01, T=1.0,
             kappa=2.0, theta=0.04, sigma=0.5, rho=-0.7, lambda_j=0.3, mu_j=-0.05, sigma_j=0.1)
    d.update(kw)
    return oracle.model(**d)


@pytest.mark.parametrize("model,scheme", MODELS)
def test_no_noise_forward_growth(oracle, model, scheme):
    """NoNoiseTest.hpp:113-204: with zero volatility every path ends at S0 e^{(r-q)T}."""
    m = M(oracle, model=model, scheme=scheme, n_steps=365, S0=100.0, v0=0.0, r=0.05, q=0.02,
          kappa=0.0, theta=0.0, sigma=0.0, rho=0.0, lambda_j=0.0, mu_j=0.0, sigma_j=0.0)
    out = oracle.simulate(m, 100)
    assert np.abs(out["S_T"] - 100.0 * math.exp(0.03)).max() < 1e-8


@pytest.mark.parametrize("model,scheme", MODELS)
def test_black_scholes_limit(oracle, model, scheme):
    """BlackScholesTest.hpp:46-107: kappa=theta=sigma=rho=0 (lambda=0) -> BS within 3 SE."""
    m = M(oracle, model=model, scheme=scheme, n_steps=16, v0=0.04, r=0.04, q=0.0, kappa=0.0,
          theta=0.0, sigma=0.0, rho=0.0, lambda_j=0.0, mu_j=0.0, sigma_j=0.0)
    s = oracle.price(m, 200_000, [oracle.Payoff(0, 0, 1.0, 1.0)], 100.0)[0]
    exact = oracle.bs_call(100.0, 100.0, 1.0, 0.04, 0.0, 0.2)
    assert abs(s.price - exact) <= 3 * s.se


@pytest.mark.parametrize("model,scheme", [(0, 2), (1, 2), (1, 0)])
def test_martingale(oracle, model, scheme):
    """E[e^{-rT} S_T] = S0 e^{-qT}: an asset-or-nothing call with a tiny strike."""
    m = M(oracle, model=model, scheme=scheme, n_steps=40)
    s = oracle.price(m, 200_000, [oracle.Payoff(9, 0, 1e-9, 1.0)], 1.0)[0]
    assert abs(s.price - math.exp(-0.01)) <= 3.5 * s.se


# Seed 42, path 16687499: word 3 of Philox block (path, step 0, slot 0) is at
# least 2^32 - 128, so the FP32 uniform map gives exactly 1.0 for the step's
# first Poisson uniform (tests/golden/find_fp32_uniform_one.py finds it).
U_ONE_SEED, U_ONE_PATH = 42, 16687499


@pytest.mark.parametrize("scheme", [0, 2])
def test_zero_jump_intensity_never_jumps(oracle, scheme):
    """MSchemes.hpp:502-509: Knuth's loop jumps while the uniform product is >=
    e^{-lambda dt}; the reference's uniforms are below 1, so lambda = 0 never
    jumps.  On the path whose first Poisson uniform is 1.0 in the FP32 map, the
    jump size parameters must not matter at lambda = 0 -- and do at a lambda
    whose threshold rounds to 1 (the planted failure a threshold of exactly 1
    would give)."""
    assert oracle.lib().qko_uniform(oracle.philox((U_ONE_PATH, 0, 0, 0), (U_ONE_SEED, 0))[3], 1) == 1.0
    base = dict(model=1, scheme=scheme, fp32_semantics=1, n_steps=4, seed=U_ONE_SEED)
    def paths(**kw):
        d = dict(base)
        d.update(kw)
        return oracle.simulate(M(oracle, **d), 1, path0=U_ONE_PATH, store=True)["paths"]
    a = paths(lambda_j=0.0, mu_j=-0.5, sigma_j=0.3)
    b = paths(lambda_j=0.0, mu_j=0.0, sigma_j=0.0)
    assert np.array_equal(a, b)
    c = paths(lambda_j=1e-300, mu_j=-0.5, sigma_j=0.3)
    d = paths(lambda_j=1e-300, mu_j=0.0, sigma_j=0.0)
    assert not np.array_equal(c, d)
    assert np.array_equal(b, d)


def test_qe_matches_heston_cf_coarse_grid(oracle):
    """Andersen QE is nearly unbiased at coarse grids (Rouah p.28 parameters, N=32)."""
    m = M(oracle, scheme=2, n_steps=32, v0=0.05, r=0.03, q=0.02, T=0.5, kappa=5.0, theta=0.05,
          sigma=0.5, rho=-0.8)
    s = oracle.price(m, 400_000, [oracle.Payoff(0, 0, 1.0, 1.0)], 100.0)[0]
    assert abs(s.price - 6.2528) <= 3 * s.se + 2e-4


def test_qe_zero_kappa_branch(oracle):
    """kappa < eps uses m = v and s^2 = sigma^2 v dt (MSchemes.hpp:322-326)."""
    m = M(oracle, scheme=2, n_steps=64, kappa=0.0, theta=0.0, v0=0.04, sigma=0.3, rho=-0.5, q=0.0)
    s = oracle.price(m, 300_000, [oracle.Payoff(0, 0, 1.0, 1.0)], 100.0)[0]
    exact = oracle.heston_call(100.0, 100.0, 1.0, 0.03, 0.0, 0.04, 0.0, 0.0, 0.3, -0.5)
    assert abs(s.price - exact) <= 3 * s.se + 0.01


def test_bates_qe_matches_cf(oracle):
    """BatesTest.hpp:52-123 parameters."""
    m = M(oracle, model=1, scheme=2, n_steps=64, v0=0.05, r=0.03, q=0.02, T=0.5, kappa=5.0,
          theta=0.05, sigma=0.5, rho=-0.8, lambda_j=0.11, mu_j=-0.15, sigma_j=0.11)
    s = oracle.price(m, 400_000, [oracle.Payoff(0, 0, 1.0, 1.0)], 100.0)[0]
    exact = oracle.bates_call(100, 100, 0.5, 0.03, 0.02, 0.05, 5.0, 0.05, 0.5, -0.8, 0.11, -0.15, 0.11)
    assert abs(s.price - exact) <= 3 * s.se + 2e-4


def brute_payoff(t, right, K, B, path, S0):
    """Payoff.hpp semantics written from the definitions on a whole path
    (path excludes S0; max/min include S0; strict vs non-strict barriers)."""
    ST = path[-1]
    mx, mn = max(S0, path.max()), min(S0, path.min())
    avg = path.mean()
    van = (lambda S, k: max(S - k, 0.0)) if right == 0 else (lambda S, k: max(k - S, 0.0))
    if t == 0:
        return van(ST, K)
    if t == 1:
        return van(avg, K)
    if t == 2:
        return van(ST, K) if mx < B else 0.0
    if t == 3:
        return van(ST, K) if mn > B else 0.0
    if t == 4:
        return van(ST, K) if mx >= B else 0.0
    if t == 5:
        return van(ST, K) if mn <= B else 0.0
    if t == 6:
        return max(ST - mn, 0.0) if right == 0 else max(mx - ST, 0.0)
    if t == 7:
        return max(mx - K, 0.0) if right == 0 else max(K - mn, 0.0)
    if t == 8:
        return float(ST > K) if right == 0 else float(ST < K)
    if t == 9:
        return ST * float(ST > K) if right == 0 else ST * float(ST < K)
    raise ValueError(t)


@pytest.mark.parametrize("model,scheme", [(0, 2), (1, 0)])
def test_payoffs_brute_force(oracle, model, scheme):
    m = M(oracle, model=model, scheme=scheme, n_steps=12)
    n = 300
    out = oracle.simulate(m, n, store=True)
    paths = out["paths"]
    for t in range(10):
        for right in (0, 1):
            K, B = (1.02, 1.1) if t not in (3, 5) else (1.02, 0.92)
            p = oracle.Payoff(t, right, K, B)
            want = np.array([brute_payoff(t, right, K, B, paths[:, i], 1.0) for i in range(n)])
            got = np.array([oracle.payoff_eval(p, out["S_T"][i], out["sum"][i], out["max"][i],
                                               out["min"][i], 12) for i in range(n)])
            assert np.allclose(got, want, rtol=1e-12, atol=1e-15), (t, right)
            s = oracle.price(m, n, [p], 100.0)[0]
            scale = 1.0 if t == 8 else 100.0
            disc = math.exp(-0.03)
            assert math.isclose(s.price, want.mean() * disc * scale, rel_tol=1e-12, abs_tol=1e-14)
            assert math.isclose(s.se, want.std(ddof=1) / math.sqrt(n) * disc * scale, rel_tol=1e-9,
                                abs_tol=1e-14)
            assert s.itm == int((want > 1e-12).sum())


def test_store_layout_and_prefix(oracle):
    m = M(oracle, model=1, scheme=2, n_steps=7)
    a = oracle.simulate(m, 50, store=True)
    b = oracle.simulate(m, 20, path0=30, store=True)
    assert np.array_equal(a["paths"][:, 30:], b["paths"])
    assert np.array_equal(a["paths"][-1], a["S_T"])


def test_chan_merge_matches_two_pass(oracle):
    rng = np.random.default_rng(0)
    x = rng.lognormal(size=1001)
    n, mean, m2 = 0.0, 0.0, 0.0
    for part in np.split(x, [3, 10, 400, 401, 900]):
        n, mean, m2 = oracle.chan_merge((n, mean, m2), (len(part), part.mean(), ((part - part.mean()) ** 2).sum()))
    assert n == 1001
    assert math.isclose(mean, x.mean(), rel_tol=1e-14)
    assert math.isclose(m2, ((x - x.mean()) ** 2).sum(), rel_tol=1e-12)


def test_laguerre_basis(oracle):
    from scipy.special import eval_laguerre
    for X in (0.0, 0.3, 0.9, 1.7):
        L = oracle.laguerre4(X)
        assert np.allclose(L, [eval_laguerre(k, X) for k in range(4)], rtol=1e-14, atol=1e-15)


def test_lsm_normal_equations(oracle):
    rng = np.random.default_rng(1)
    X = rng.uniform(0.5, 1.2, size=500)
    Y = rng.uniform(0, 0.3, size=500)
    L = np.stack([oracle.laguerre4(x) for x in X])
    sums = np.concatenate([[len(X)], (L.T @ L).ravel(), L.T @ Y])
    C, fb = oracle.lsm_solve(sums)
    assert not fb
    ref = np.linalg.solve(L.T @ L + 1e-8 * np.eye(4), L.T @ Y)
    assert np.allclose(C, ref, rtol=1e-7, atol=1e-10)
    ls = np.linalg.lstsq(L, Y, rcond=None)[0]
    assert np.allclose(C, ls, rtol=1e-3, atol=1e-5)   # the 1e-8 ridge on a badly conditioned Gram
    # fewer than 4 ITM paths: constant fallback B0 / count (LSMAmerican.hpp:605-613)
    C, fb = oracle.lsm_solve(np.concatenate([[3.0], np.zeros(16), [0.6, 1, 1, 1]]))
    assert fb and np.allclose(C, [0.2, 0, 0, 0])
    # rank-one Gram: the 1e-8 ridge keeps every pivot above 1e-14, and the
    # ridge solution reproduces the data at the single abscissa
    L1 = np.stack([oracle.laguerre4(0.8)] * 10)
    C, fb = oracle.lsm_solve(np.concatenate([[10.0], (L1.T @ L1).ravel(), L1.T @ np.full(10, 0.2)]))
    assert not fb and abs(C @ L1[0] - 0.2) < 1e-6
    # an indefinite matrix (negative pivot) takes the fallback
    bad = np.concatenate([[10.0], -np.eye(4).ravel(), [2.0, 0, 0, 0]])
    C, fb = oracle.lsm_solve(bad)
    assert fb and np.allclose(C, [0.2, 0, 0, 0])


IKONEN_TOIVANEN = {8.0: 2.0, 9.0: 1.107641, 10.0: 0.520030, 11.0: 0.213668, 12.0: 0.082036}


@pytest.mark.parametrize("S0", sorted(IKONEN_TOIVANEN))
def test_american_put_ikonen_toivanen(oracle, S0):
    """AmericanOption.hpp:109-177 (Rouah p.212): Heston implicit Milstein LSM put,
    tolerance 3 SE + max(2% ref, 0.005).  Reduced here to 100k paths, dt=T/250."""
    K = 10.0
    m = M(oracle, scheme=1, n_steps=250, seed=3, S0=1.0, v0=0.0625, r=0.1, q=0.0, T=0.25, kappa=5.0,
          theta=0.16, sigma=0.9, rho=0.1)
    s = oracle.price_american(m, 100_000, 1, K / S0, S0)
    ref = IKONEN_TOIVANEN[S0]
    assert abs(s.price - ref) <= 3 * s.se + max(0.02 * ref, 0.005), (s.price, ref, s.se)


def test_greek_put_call_parity(oracle):
    """PutCallParityTest.hpp:194-277 identities under common random numbers."""
    m = M(oracle, scheme=1, n_steps=100, v0=0.04, r=0.05, q=0.0, kappa=2.0, theta=0.04, sigma=0.3,
          rho=-0.7)
    b = oracle.Bumps(1, 1, 1, 1, 1, 0, 0.01, 0.02, 0.001, 0.02)
    n = 100_000
    gc = oracle.greeks(m, n, oracle.Payoff(0, 0, 1.0, 1.0), b, 100.0)
    gp = oracle.greeks(m, n, oracle.Payoff(0, 1, 1.0, 1.0), b, 100.0)
    K, T, r = 100.0, 1.0, 0.05

    def tol(scale, mult):
        return max(0.01 * max(abs(scale), 1.0), mult * gc.base_se, 1e-4)

    assert abs((gc.delta - gp.delta) - 1.0) <= tol(1.0, 3.3)
    assert abs(gc.gamma - gp.gamma) <= tol(1.0, 3.3)
    # Per path C - P = e^{-rT}(S_T - K), so vega(C) - vega(P) is the bumped
    # difference of mean S_T, zero in expectation; its noise is that of the
    # per-path difference of S_T between the two variance bumps (common
    # random numbers), divided by the bump width -- far above the price SE.
    vol0 = math.sqrt(m.v0)
    up, dn = M(oracle, scheme=1, n_steps=100, v0=(vol0 + 0.02) ** 2, r=0.05, q=0.0, kappa=2.0, theta=0.04,
               sigma=0.3, rho=-0.7), M(oracle, scheme=1, n_steps=100, v0=(vol0 - 0.02) ** 2, r=0.05, q=0.0,
                                       kappa=2.0, theta=0.04, sigma=0.3, rho=-0.7)
    d = oracle.simulate(up, n)["S_T"] - oracle.simulate(dn, n)["S_T"]
    se_vega_diff = 100.0 * math.exp(-r * T) * d.std() / math.sqrt(n) / (2 * 0.02)
    assert abs(gc.vega - gp.vega) <= 4.0 * se_vega_diff + 1e-9, (gc.vega - gp.vega, se_vega_diff)
    # vomma parity is noise-dominated at this path count; the GPU suite checks it
    # at the reference's own size (1M paths, 365 steps)
    assert abs(gc.vanna - gp.vanna) <= tol(1.0, 3.3)
    exp_rho = K * T * math.exp(-r * T)
    assert abs((gc.rho - gp.rho) - exp_rho) <= tol(exp_rho, 3.3)
    exp_theta = -r * K * math.exp(-r * T)
    assert abs((gc.theta - gp.theta) - exp_theta) <= tol(exp_theta, 5.0)
    assert gc.n_pricings == 12


def test_basis_orders_and_weight(oracle):
    """LaguerreP02/P03/P04 = L0..L2, L0..L3, L0..L4 (LSMAmerican.hpp:172-198) against
    scipy's Laguerre polynomials; the weighted variant is exp(-X/2) times them."""
    from scipy.special import eval_laguerre
    for X in (0.0, 0.37, 0.9, 1.0, 1.6, 3.2):
        for nb in (3, 4, 5):
            L = oracle.basis(X, nb)
            assert np.allclose(L, [eval_laguerre(k, X) for k in range(nb)], rtol=1e-13, atol=1e-14)
            W = oracle.basis(X, nb, weighted=True)
            assert np.allclose(W, np.exp(-X / 2) * L, rtol=1e-15, atol=1e-16)


@pytest.mark.parametrize("nb", [3, 5])
def test_lsm_solve_other_orders(oracle, nb):
    """The ridge normal equations (M + 1e-8 I) C = B for 3 and 5 basis functions
    against numpy, and the count < nb fallback (LSMAmerican.hpp:516-613)."""
    rng = np.random.default_rng(nb)
    X = rng.uniform(0.5, 1.0, 400)
    L = np.stack([oracle.basis(x, nb) for x in X])
    Y = np.maximum(1.0 - X, 0) + 0.05 * rng.standard_normal(400)
    M = L.T @ L
    C, fb = oracle.lsm_solve_n(np.concatenate([[400.0], M.ravel(), L.T @ Y]), nb)
    ref = np.linalg.solve(M + 1e-8 * np.eye(nb), L.T @ Y)
    assert not fb and np.allclose(L @ C, L @ ref, rtol=1e-7, atol=1e-9)
    C, fb = oracle.lsm_solve_n(np.concatenate([[nb - 1.0], np.zeros(nb * nb), [0.3] + [1.0] * (nb - 1)]), nb)
    assert fb and C[0] == pytest.approx(0.3 / (nb - 1)) and not C[1:].any()


@pytest.mark.parametrize("nb,weighted", [(3, False), (5, False), (4, True)])
def test_american_other_bases_ikonen_toivanen(oracle, nb, weighted):
    """The other regression bases on the S0 = K = 10 case of the Ikonen-Toivanen table
    (0.520030), same tolerance rule as the reference's test."""
    m = M(oracle, scheme=1, n_steps=250, seed=3, S0=1.0, v0=0.0625, r=0.1, q=0.0, T=0.25, kappa=5.0,
          theta=0.16, sigma=0.9, rho=0.1)
    s = oracle.price_american(m, 100_000, 1, 1.0, 10.0, n_basis=nb, weighted=weighted)
    assert abs(s.price - 0.520030) <= 3 * s.se + max(0.02 * 0.520030, 0.005), (s.price, s.se)


def test_price_windows_merge_to_the_whole(oracle):
    """Pricing path windows [0, a), [a, b), [b, n) of the stream and merging their
    moments (Chan, OptionPricer.hpp:152-172) gives the one-shot price: the
    stream is keyed by the global path index."""
    m = M(oracle, model=0, scheme=2, n_steps=30, seed=42, S0=1.0, v0=0.04, r=0.03, q=0.0, T=1.0, kappa=2.0,
          theta=0.04, sigma=0.5, rho=-0.7)
    pays = [oracle.Payoff(1, 0, 1.0, 1.0), oracle.Payoff(2, 0, 1.0, 1.3), oracle.Payoff(6, 1, 1.0, 1.0)]
    n, cuts = 20_000, (0, 6_401, 13_000, 20_000)
    whole = oracle.price(m, n, pays, 100.0)
    parts = [oracle.price_window(m, a, b - a, pays, 100.0) for a, b in zip(cuts, cuts[1:])]
    for j, w in enumerate(whole):
        mg = oracle.merge_stats([p[j] for p in parts], math.exp(-m.r * m.T), 100.0)
        assert mg.n == w.n and mg.itm == w.itm
        assert math.isclose(mg.price, w.price, rel_tol=1e-12)
        assert math.isclose(mg.se, w.se, rel_tol=1e-9)
