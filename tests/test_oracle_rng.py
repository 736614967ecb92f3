"""Pins the oracle's random stream: Philox4x32-10 known-answer vectors, the
uniform maps, and the reference's normal-moment test."""
import numpy as np


# Known-answer vectors of Philox4x32-10 published with Random123 (kat_vectors):
# (counter, key) -> output.
KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


def test_philox_known_answers(oracle):
    for ctr, key, want in KAT:
        assert tuple(oracle.philox(ctr, key)) == want


def test_uniform_maps(oracle):
    u = oracle.lib().qko_uniform
    # FP64 map: open interval (0, 1), centred cells of width 2^-32
    assert u(0, 0) == 2.0 ** -33
    assert u(0xFFFFFFFF, 0) == 1.0 - 2.0 ** -33
    # FP32 map: (0, 1]; the top words round up to 1
    assert u(0, 1) == 2.0 ** -33
    assert u(0xFFFFFFFF, 1) == 1.0
    # below 2^23 every value is exact in FP32, so the two maps agree
    for w in (1, 2, 3, 12345, (1 << 23) - 1):
        assert u(w, 1) == u(w, 0)
    # round to nearest even at 2^28 (FP32 spacing 32): ties go to the even multiple
    assert u(0x1234560F, 1) == u(0x12345600, 1) == u(0x12345610, 1)
    assert u(0x12345611, 1) == u(0x12345620, 1) > u(0x12345600, 1)
    # against numpy's float32 arithmetic, and within two roundings (half an FP32
    # ulp of 1 each) of the exact map
    rng = np.random.default_rng(3)
    ws = np.concatenate([rng.integers(0, 2 ** 32, 2000, dtype=np.uint64),
                         np.array([2 ** 24 - 1, 2 ** 24, 2 ** 24 + 1, 2 ** 31, 2 ** 32 - 129, 2 ** 32 - 128],
                                  dtype=np.uint64)])
    got = np.array([u(int(w), 1) for w in ws])
    want = (ws.astype(np.uint32).astype(np.float32) * np.float32(2.0 ** -32) + np.float32(2.0 ** -33))
    assert np.array_equal(got, want.astype(np.float64))
    exact = (ws.astype(np.float64) + 0.5) * 2.0 ** -32
    assert np.max(np.abs(got - exact)) <= 2.0 ** -24
    order = np.argsort(ws, kind="stable")
    assert np.all(np.diff(got[order]) >= 0) and got.min() > 0 and got.max() <= 1.0


def test_normal_moments(oracle):
    """GaussianRNGenTest.hpp:200-208: |mean| < 0.005, |var - 1| < 0.01."""
    for u24 in (False, True):
        z = oracle.normals(184467440737095, 20_000, 50, u24=u24)
        for k in (0, 1):
            x = z[:, :, k].ravel()
            assert abs(x.mean()) < 0.005
            assert abs(x.var() - 1.0) < 0.01
        # the two Box-Muller outputs are uncorrelated, as are consecutive steps
        assert abs(np.corrcoef(z[:, :, 0].ravel(), z[:, :, 1].ravel())[0, 1]) < 0.005
        assert abs(np.corrcoef(z[:, :-1, 0].ravel(), z[:, 1:, 0].ravel())[0, 1]) < 0.005
        # fourth moment of a standard normal is 3
        assert abs((z[:, :, 0] ** 4).mean() - 3.0) < 0.05


def test_streams_depend_on_seed_path_and_step(oracle):
    a = oracle.normals(1, 4, 3)
    b = oracle.normals(2, 4, 3)
    assert not np.allclose(a, b)
    assert len(np.unique(a[:, :, 0])) == 12
