"""Black-Scholes closed forms used as known answers (test helper).

With kappa = theta = sigma = rho = lambda = 0 the Heston / Bates models keep
v = v0 on every step, so the simulated spot is geometric Brownian motion at
volatility sqrt(v0): the reference's own Black-Scholes test uses that limit
(BlackScholesTest.hpp:46-83, BlackScholesExact.hpp:46-65).  The Greeks below
are the textbook derivatives of the Black-Scholes prices with continuous
dividend yield q; theta is -dV/dT (value lost as maturity shortens), which is
what the reference's one-sided stencil (P(T - h) - P(T)) / h estimates
(MCEngine.hpp:651-664); vega, vomma and vanna differentiate in the
volatility sqrt(v0), which is what the reference bumps (MCEngine.hpp:614-631).
"""
import math

from scipy.stats import norm


def _d(S, K, T, r, q, vol):
    sq = math.sqrt(T)
    d1 = (math.log(S / K) + (r - q + 0.5 * vol * vol) * T) / (vol * sq)
    return d1, d1 - vol * sq


def call_price(S, K, T, r, q, vol):
    d1, d2 = _d(S, K, T, r, q, vol)
    return S * math.exp(-q * T) * norm.cdf(d1) - K * math.exp(-r * T) * norm.cdf(d2)


def call_greeks(S, K, T, r, q, vol):
    d1, d2 = _d(S, K, T, r, q, vol)
    sq = math.sqrt(T)
    eq, er = math.exp(-q * T), math.exp(-r * T)
    pd1 = norm.pdf(d1)
    vega = S * eq * pd1 * sq
    return {
        "price": call_price(S, K, T, r, q, vol),
        "delta": eq * norm.cdf(d1),
        "gamma": eq * pd1 / (S * vol * sq),
        "vega": vega,
        "vomma": vega * d1 * d2 / vol,
        "vanna": -eq * pd1 * d2 / vol,
        "rho": K * T * er * norm.cdf(d2),
        "theta": -S * eq * pd1 * vol / (2 * sq) - r * K * er * norm.cdf(d2) + q * S * eq * norm.cdf(d1),
    }


def cash_binary_call_greeks(S, K, T, r, q, vol):
    """Cash-or-nothing call paying 1: price e^{-rT} N(d2)."""
    d1, d2 = _d(S, K, T, r, q, vol)
    sq = math.sqrt(T)
    er = math.exp(-r * T)
    pd2 = norm.pdf(d2)
    dd2_dT = -math.log(S / K) / (2 * vol * T * sq) + (r - q - 0.5 * vol * vol) / (2 * vol * sq)
    return {
        "price": er * norm.cdf(d2),
        "delta": er * pd2 / (S * vol * sq),
        "gamma": -er * pd2 * d1 / (S * S * vol * vol * T),
        "vega": -er * pd2 * d1 / vol,
        "rho": -T * er * norm.cdf(d2) + er * pd2 * sq / vol,
        "theta": r * er * norm.cdf(d2) - er * pd2 * dd2_dT,
    }
