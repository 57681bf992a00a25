"""Launch-shape tuner, host logic (no GPU): candidates, choice, cache file."""

import json

from magic_amd import tune


def test_candidates_are_powers_of_two_below_the_plan_block():
    assert tune.grid_candidates(1) == [1]
    assert tune.grid_candidates(4) == [1, 2, 4]
    assert tune.grid_candidates(6) == [1, 2, 4, 6]
    assert tune.grid_candidates(16) == [1, 2, 4, 8, 16]


def test_choice_is_the_fastest_and_ties_keep_the_plan_block():
    assert tune.choose({1: 2.9, 2: 3.2, 4: 3.6}, 4) == 1
    assert tune.choose({1: 3.6, 2: 3.2, 4: 3.6}, 4) == 2
    assert tune.choose({1: 3.0, 2: 3.2, 4: 3.0}, 4) == 0   # tie: the plan's own (rb = 0)
    assert tune.choose({4: 1.0}, 4) == 0


def test_cache_file_round_trip(tmp_path, monkeypatch):
    path = tmp_path / "sub" / "tune.json"
    monkeypatch.setenv("MAGIC_AMD_TUNE_CACHE", str(path))
    tune._write_file("k1", dict(rb=1, ms={"1": 2.9}))
    tune._write_file("k2", dict(rb=0, ms={"4": 3.0}))
    data = json.loads(path.read_text())
    assert data["k1"]["rb"] == 1 and data["k2"]["rb"] == 0
    assert tune._read_file() == data
    path.write_text("not json")
    assert tune._read_file() == {}
