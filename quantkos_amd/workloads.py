"""Synthetic request streams for the serving path.

The config-4 stream: pricing requests with Poisson arrivals at a fixed rate.
  * contract mix: European 30 %, Asian 20 %, barrier 15 % (up/down x in/out),
    lookback 10 % (fixed/floating), binary 5 % (cash/asset), American LSM 20 %;
    call/put 50/50;
  * model Heston or Bates 50/50; Bates jumps lambda=0.1, muJ=-0.05, sigmaJ=0.10;
  * Andersen QE scheme; paths log-uniform in [1e4, 1e7]; steps uniform in
    {12, 52, 252, 504};
  * parameters uniform: kappa [0.5,5], theta [0.01,0.09], sigma [0.1,1.0],
    rho [-0.9,0], v0 [0.01,0.09], K/S0 [0.8,1.2], T [0.1,2]; r=0.03, q=0;
    S0=100 (normalised to 1); up barriers at 1.3 S0, down barriers at 0.75 S0;
  * request i uses seed 1000 + i.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .engine import contract, payoff

MIX = [(_lib.EUROPEAN, 0.30), (_lib.ASIAN, 0.20), ("barrier", 0.15), ("lookback", 0.10),
       ("binary", 0.05), (_lib.AMERICAN, 0.20)]
STEPS = (12, 52, 252, 504)


def config4_stream(n=4096, seed=2024, precision=_lib.FP32, paths_range=(1e4, 1e7), rate=None,
                   scheme=_lib.QE):
    """Returns (contracts, payoffs, arrival_s).  `rate` (requests/s) draws
    Poisson arrivals; None puts every arrival at 0 (a burst)."""
    rng = np.random.default_rng(seed)
    kinds = rng.choice(len(MIX), size=n, p=[w for _, w in MIX])
    contracts, payoffs = [], []
    lo, hi = np.log(paths_range[0]), np.log(paths_range[1])
    for i in range(n):
        kind = MIX[kinds[i]][0]
        right = int(rng.integers(0, 2))
        if kind == "barrier":
            t = int(rng.choice([_lib.BARRIER_UP_OUT, _lib.BARRIER_DOWN_OUT, _lib.BARRIER_UP_IN,
                                _lib.BARRIER_DOWN_IN]))
        elif kind == "lookback":
            t = int(rng.choice([_lib.LOOKBACK_FIXED, _lib.LOOKBACK_FLOATING]))
        elif kind == "binary":
            t = int(rng.choice([_lib.BINARY_CASH, _lib.BINARY_ASSET]))
        else:
            t = int(kind)
        model = int(rng.integers(0, 2))
        n_paths = int(np.exp(rng.uniform(lo, hi)))
        n_steps = int(rng.choice(STEPS))
        kappa = rng.uniform(0.5, 5.0)
        theta = rng.uniform(0.01, 0.09)
        sigma = rng.uniform(0.1, 1.0)
        rho = rng.uniform(-0.9, 0.0)
        v0 = rng.uniform(0.01, 0.09)
        K = rng.uniform(0.8, 1.2)
        T = rng.uniform(0.1, 2.0)
        B = 1.3 if t in (_lib.BARRIER_UP_OUT, _lib.BARRIER_UP_IN) else 0.75
        contracts.append(contract(model=model, scheme=scheme, precision=precision, n_steps=n_steps,
                                  n_paths=n_paths, seed=1000 + i, S0=1.0, v0=v0, r=0.03, q=0.0, T=T,
                                  kappa=kappa, theta=theta, sigma=sigma, rho=rho, lambda_j=0.1,
                                  mu_j=-0.05, sigma_j=0.10, price_scale=100.0))
        payoffs.append(payoff(t, right, K, B))
    if rate is None:
        arrivals = np.zeros(n)
    else:
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n))
        arrivals -= arrivals[0]
    return contracts, payoffs, list(arrivals)


def path_steps(contracts) -> int:
    return int(sum(c.n_paths * c.n_steps for c in contracts))
