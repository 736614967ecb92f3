"""Launch schedules change speed only: prices under every lanes-per-thread
choice agree to rounding, and a (quick) autotuner run writes a table that the
engine applies."""
import math

import pytest

import quantkos_amd as qk
from quantkos_amd import autotune

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [qk.FP64, qk.FP32])
@pytest.mark.parametrize("n_lanes", [3, 13])
def test_lane_schedules_agree(engine, precision, n_lanes):
    lanes = autotune._lane_contracts(precision, qk.MILSTEIN, n_lanes, 20_000, 40)
    pays = [qk.payoff(qk.EUROPEAN, qk.CALL, 1.0), qk.payoff(qk.BARRIER_UP_OUT, qk.CALL, 1.0, 1.2)]
    res = {}
    for lpt in (1, 2, 4):
        engine.set_schedule([(precision, qk.MILSTEIN, n_lanes, lpt)])
        assert engine.lanes_per_thread(precision, qk.MILSTEIN, n_lanes) == lpt
        res[lpt] = engine.price_lanes(lanes, pays)
    if engine.tuning:
        autotune.apply(engine, engine.tuning)
    else:
        engine.set_schedule([])
    tol = 1e-12 if precision == qk.FP64 else 2e-6
    for lpt in (2, 4):
        for row_a, row_b in zip(res[1], res[lpt]):
            for a, b in zip(row_a, row_b):
                assert math.isclose(a.price, b.price, rel_tol=tol, abs_tol=1e-12), (lpt, a.price, b.price)
                assert a.itm == b.itm


def test_quick_autotune_writes_and_applies_a_table(engine, tmp_path):
    out = str(tmp_path / "t.json")
    table = autotune.autotune(engine, out, quick=True, log=lambda m: None)
    assert autotune.load(out) == table
    e = qk.Engine(0, tuning=out)
    try:
        for ent in table["lanes"]:
            assert e.lanes_per_thread(ent["precision"], ent["scheme"], ent["n_lanes"]) == ent["lanes_per_thread"]
        with qk.Server(e) as srv:
            c = qk.contract(scheme=qk.QE, n_paths=10_000, n_steps=12)
            r = srv.wait(srv.submit(c, qk.payoff(qk.EUROPEAN, qk.CALL, 1.0)))
            assert r.status == 0 and r.stats.n == 10_000
    finally:
        e.close()
