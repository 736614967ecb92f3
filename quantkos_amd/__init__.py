"""MI355X-native Monte Carlo hot path of QuantKos: gfx950 HIP kernels behind a C ABI
(include/qk.h, libquantkos.so) and a C++20 host that keeps QuantKos's YAML schema."""
from . import _lib
from ._lib import (AMERICAN, ASIAN, BARRIER_DOWN_IN, BARRIER_DOWN_OUT, BARRIER_UP_IN,
                   BARRIER_UP_OUT, BATES, BINARY_ASSET, BINARY_CASH, CALL, EULER, EUROPEAN, FP32,
                   FP64, HESTON, LAGUERRE_P02, LAGUERRE_P03, LAGUERRE_P04, LOOKBACK_FIXED,
                   LOOKBACK_FLOATING, MILSTEIN, PUT, QE, WINDOW_CHUNK_PATHS, Moments)
from .engine import Engine, QKError, Server, analytic_european, bumps, contract, payoff

__all__ = [n for n in dir() if not n.startswith("_")]
