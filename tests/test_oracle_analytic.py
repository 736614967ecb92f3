"""Pins the oracle's closed forms: Gauss-Legendre rule, Black-Scholes, Heston
and Bates characteristic-function prices, against known answers and
independent computations (numpy / scipy quadrature, Merton's series)."""
import math

import numpy as np
from scipy import integrate, stats


def test_gl64_matches_numpy_and_tables(oracle):
    x, w = oracle.gl64()
    X, W = np.polynomial.legendre.leggauss(64)
    assert np.abs(x - X).max() < 1e-14
    assert np.abs(w - W).max() < 1e-14
    # tabulated values (Integration.hpp:43-68): innermost and outermost nodes
    assert abs(x[32] - 0.0243502926634244) < 1e-15 and abs(w[32] - 0.0486909570091397) < 1e-15
    assert abs(x[63] - 0.9993050417357721) < 1e-15 and abs(w[63] - 0.0017832807216964) < 1e-15
    assert abs(w.sum() - 2.0) < 1e-13


def test_black_scholes_against_quadrature(oracle):
    S0, K, T, r, q, vol = 100.0, 95.0, 0.7, 0.04, 0.01, 0.25

    def integrand(z):
        ST = S0 * math.exp((r - q - 0.5 * vol * vol) * T + vol * math.sqrt(T) * z)
        return max(ST - K, 0.0) * stats.norm.pdf(z)

    val, _ = integrate.quad(integrand, -12, 12, limit=200, points=[0.0])
    assert abs(oracle.bs_call(S0, K, T, r, q, vol) - math.exp(-r * T) * val) < 1e-9


def test_heston_rouah_known_answer(oracle):
    """HestonTest.hpp:51: Rouah p.28 gives 6.2528."""
    p = oracle.heston_call(100, 100, 0.5, 0.03, 0.02, 0.05, 5.0, 0.05, 0.5, -0.8)
    assert abs(p - 6.2528) < 2e-4


def test_heston_config0_value(oracle):
    """Config-0 parameters: an independent quadrature gives 8.92941."""
    p = oracle.heston_call(100, 100, 1.0, 0.03, 0.0, 0.04, 2.0, 0.04, 0.5, -0.7)
    assert abs(p - 8.92941) < 1e-4


def test_heston_reduces_to_black_scholes(oracle):
    # negligible vol-of-vol with v0 = theta: constant variance 0.04
    p = oracle.heston_call(100, 105, 1.0, 0.03, 0.01, 0.04, 1.5, 0.04, 1e-3, -0.5)
    bs = oracle.bs_call(100, 105, 1.0, 0.03, 0.01, 0.2)
    assert abs(p - bs) < 2e-3


def test_heston_against_direct_fourier_inversion(oracle):
    """Same characteristic function, integrated adaptively to 500 with scipy:
    the oracle's single 64-point panel on [0, 100] must agree (HestonExact.hpp:145-157)."""
    S0, K, T, r, q, v0, k, th, s, rho = 100, 110, 0.75, 0.02, 0.0, 0.06, 3.0, 0.05, 0.4, -0.6

    def f(phi, j):
        u = 0.5 if j == 1 else -0.5
        b = k - rho * s if j == 1 else k
        d = np.sqrt((rho * s * 1j * phi - b) ** 2 - s * s * (2 * u * 1j * phi - phi * phi))
        g = (b - rho * s * 1j * phi - d) / (b - rho * s * 1j * phi + d)
        e = np.exp(-d * T)
        C = (r - q) * 1j * phi * T + k * th / s ** 2 * ((b - rho * s * 1j * phi - d) * T - 2 * np.log((1 - g * e) / (1 - g)))
        D = (b - rho * s * 1j * phi - d) / s ** 2 * (1 - e) / (1 - g * e)
        return np.exp(C + D * v0 + 1j * phi * np.log(S0))

    P = [0.5 + integrate.quad(lambda x: (np.exp(-1j * x * np.log(K)) * f(x, j) / (1j * x)).real,
                              1e-9, 500, limit=500)[0] / math.pi for j in (1, 2)]
    direct = S0 * math.exp(-q * T) * P[0] - K * math.exp(-r * T) * P[1]
    assert abs(oracle.heston_call(S0, K, T, r, q, v0, k, th, s, rho) - direct) < 1e-6


def test_bates_without_jumps_is_heston(oracle):
    h = oracle.heston_call(100, 100, 0.5, 0.03, 0.02, 0.05, 5.0, 0.05, 0.5, -0.8)
    b = oracle.bates_call(100, 100, 0.5, 0.03, 0.02, 0.05, 5.0, 0.05, 0.5, -0.8, 0.0, 0.0, 0.0)
    assert abs(h - b) < 1e-6


def test_bates_against_merton_series(oracle):
    """With negligible vol-of-vol the Bates price is Merton's jump diffusion:
    sum_n Pois(n; lam' T) BS(S0, K, T, r_n, sigma_n)."""
    S0, K, T, r, q, v, lam, mu, sj = 100.0, 100.0, 0.5, 0.03, 0.02, 0.05, 0.11, -0.15, 0.11
    kbar = math.exp(mu + 0.5 * sj * sj) - 1.0
    lp = lam * (1 + kbar)
    merton = 0.0
    for n in range(40):
        rn = r - lam * kbar + n * math.log(1 + kbar) / T
        sn = math.sqrt(v + n * sj * sj / T)
        w = math.exp(-lp * T) * (lp * T) ** n / math.factorial(n)
        merton += w * oracle.bs_call(S0, K, T, rn, q, sn)
    b = oracle.bates_call(S0, K, T, r, q, v, 2.0, v, 1e-3, 0.0, lam, mu, sj)
    assert abs(b - merton) < 2e-3
