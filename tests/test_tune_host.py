"""Launch-shape tuner, host logic (no GPU): candidates, choice, cache file."""

import json

from magic_amd import tune


def test_candidates_are_powers_of_two_below_the_plan_block():
    assert tune.grid_candidates(1) == [1]
    assert tune.grid_candidates(4) == [1, 2, 4]
    assert tune.grid_candidates(6) == [1, 2, 4, 6]
    assert tune.grid_candidates(16) == [1, 2, 4, 8, 16]


def test_choice_is_the_fastest_and_ties_keep_the_default():
    assert tune.fastest({1: 2.9, 2: 3.2, 4: 3.6}, 4) == 1
    assert tune.fastest({1: 3.6, 2: 3.2, 4: 3.6}, 4) == 2
    assert tune.fastest({1: 3.0, 2: 3.2, 4: 3.0}, 4) == 4   # tie: the default
    assert tune.fastest({4: 1.0, 6: 1.0}, 4) == 4


def test_cached_choice_is_not_timed_again(tmp_path, monkeypatch):
    monkeypatch.setenv("MAGIC_AMD_TUNE_CACHE", str(tmp_path / "t.json"))
    tune._write_file("shape", dict(choice=6, ms={"4": 1.2, "6": 1.0}))
    calls = []
    assert tune.timed_choice("shape", [4, 6], calls.append, 4) == 6
    assert calls == [] and tune._MEM.pop("shape") == 6
    # a stale entry (not a candidate any more) is ignored: one candidate, no timing
    assert tune.timed_choice("shape", [4], calls.append, 4) == 4
    assert calls == [] and tune._MEM.pop("shape") == 4


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


def test_tuple_choices_survive_the_json_cache(tmp_path, monkeypatch):
    monkeypatch.setenv("MAGIC_AMD_TUNE_CACHE", str(tmp_path / "t.json"))
    tune._write_file("syn", dict(choice=[4, 2], ms={}))
    calls = []
    assert tune.timed_choice("syn", [(4, 1), (6, 1), (4, 2)], calls.append, (4, 1)) == (4, 2)
    assert calls == [] and tune._MEM.pop("syn") == (4, 2)
