"""Finds a (seed, path, step) whose Bates Poisson word (word 3 of Philox block
(path, step, 0)) maps to exactly 1.0 under the FP32 uniform map, i.e. is at
least 2^32 - 128.  Used by the lambda = 0 tests: the reference's uniforms stay
below 1, so a zero jump intensity never jumps, and the threshold used for
lambda = 0 must stay above a uniform of 1.

    python tests/golden/find_fp32_uniform_one.py      # prints seed 42, path 16687499, step 0

Philox4x32-10 is restated here with numpy (vectorised over paths) and checked
against the oracle's implementation before the search.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

M0, M1, W0, W1 = 0xD2511F53, 0xCD9E8D57, 0x9E3779B9, 0xBB67AE85
MASK = np.uint64(0xFFFFFFFF)


def philox_np(c0, c1, c2, c3, k0, k1):
    c0, c1, c2, c3 = [np.asarray(x, dtype=np.uint64) for x in (c0, c1, c2, c3)]
    k0, k1 = np.uint64(k0), np.uint64(k1)
    for _ in range(10):
        p0, p1 = np.uint64(M0) * c0, np.uint64(M1) * c2
        c0, c1, c2, c3 = (p1 >> np.uint64(32)) ^ c1 ^ k0, p1 & MASK, (p0 >> np.uint64(32)) ^ c3 ^ k1, p0 & MASK
        k0, k1 = (k0 + np.uint64(W0)) & MASK, (k1 + np.uint64(W1)) & MASK
    return c0, c1, c2, c3


def main(seed=42, step=0):
    from oracle import oracle as O
    for ctr, key in [((1, 2, 3, 0), (42, 0)), ((12345, 0, 7, 0), (9, 1))]:
        got = [int(x[0]) for x in philox_np(*[np.array([v]) for v in ctr], *key)]
        assert got == list(O.philox(ctr, key)), (got, ctr, key)
    chunk = 1 << 24
    for base in range(0, 1 << 32, chunk):
        p = np.arange(base, base + chunk, dtype=np.uint64)
        w3 = philox_np(p, np.zeros_like(p), np.full_like(p, step), np.zeros_like(p), seed & 0xFFFFFFFF, seed >> 32)[3]
        hit = np.nonzero(w3 >= np.uint64((1 << 32) - 128))[0]
        if len(hit):
            path = int(base + hit[0])
            assert O.lib().qko_uniform(O.philox((path & 0xFFFFFFFF, path >> 32, step, 0), (seed & 0xFFFFFFFF, seed >> 32))[3], 1) == 1.0
            print("seed %d, path %d, step %d" % (seed, path, step))
            return seed, path, step
    raise RuntimeError("none found")


if __name__ == "__main__":
    main()
