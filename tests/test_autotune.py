"""The autotuner's host logic: table validation, candidate choice, the default
lane schedule the C ABI falls back to (no GPU needed)."""
import json

import pytest

from quantkos_amd import _lib, autotune


def test_pick_best_takes_the_fastest_and_the_first_on_ties():
    assert autotune.pick_best({1: 3.0, 2: 1.5, 4: 2.0}) == 2
    assert autotune.pick_best({4: 1.0, 1: 1.0}) == 4


def test_validate_rejects_bad_entries():
    good = {"lanes": [{"precision": 0, "scheme": 2, "n_lanes": 13, "lanes_per_thread": 4}],
            "server": {"item_path_steps": 131072, "blocks_per_cu": 8, "max_inflight": 3}}
    autotune.validate(good)
    bad = json.loads(json.dumps(good))
    bad["lanes"][0]["lanes_per_thread"] = 3
    with pytest.raises(ValueError):
        autotune.validate(bad)
    bad = json.loads(json.dumps(good))
    bad["lanes"][0]["n_lanes"] = 65
    with pytest.raises(ValueError):
        autotune.validate(bad)
    bad = json.loads(json.dumps(good))
    bad["server"]["blocks_per_cu"] = -1
    with pytest.raises(ValueError):
        autotune.validate(bad)
    bad = json.loads(json.dumps(good))
    bad["server"]["overlap_min_backlog"] = -2
    with pytest.raises(ValueError):
        autotune.validate(bad)


def test_load_round_trip_and_missing_file(tmp_path):
    assert autotune.load(str(tmp_path / "none.json")) is None
    t = {"lanes": [{"precision": 1, "scheme": 1, "n_lanes": 12, "lanes_per_thread": 2}], "server": {}}
    p = tmp_path / "t.json"
    p.write_text(json.dumps(t))
    assert autotune.load(str(p)) == t


def test_committed_table_is_valid():
    t = autotune.load()
    if t is None:
        pytest.skip("no tuned table for this architecture")
    assert t["lanes"] and set(t["server"]) >= {"item_path_steps", "blocks_per_cu", "max_inflight"}


def test_default_lane_schedule_without_a_table():
    L = _lib.load()
    assert L.qk_lanes_per_thread(None, _lib.FP32, _lib.QE, 13) == 4
    assert L.qk_lanes_per_thread(None, _lib.FP64, _lib.EULER, 3) == 1
    assert L.qk_set_schedule(None, None, 0) == -1


def test_pick_server_prefers_the_lower_p99_within_the_throughput_band():
    a, b, c = ("a",), ("b",), ("c",)
    rates = {a: 100.0, b: 98.0, c: 90.0}
    # b is within 3 % of the best and has the lower p99 (though a higher p50);
    # c has the lowest latencies but is outside the band
    assert autotune.pick_server(rates, {a: (9.0, 230.0), b: (20.0, 200.0), c: (1.0, 10.0)}) == b
    # p99 within 5 % of the smallest: the higher throughput decides
    assert autotune.pick_server(rates, {a: (20.0, 208.0), b: (9.0, 200.0)}) == a
    # without latency figures the fastest wins
    assert autotune.pick_server(rates, {}) == a
    # a tighter band excludes b
    assert autotune.pick_server(rates, {a: (9.0, 230.0), b: (20.0, 200.0)}, tolerance=0.01) == a


def test_eight_lanes_per_thread_only_in_fp32():
    ok = {"lanes": [{"precision": _lib.FP32, "scheme": 1, "n_lanes": 13, "lanes_per_thread": 8}], "server": {}}
    autotune.validate(ok)
    bad = {"lanes": [{"precision": _lib.FP64, "scheme": 1, "n_lanes": 13, "lanes_per_thread": 8}], "server": {}}
    with pytest.raises(ValueError):
        autotune.validate(bad)
