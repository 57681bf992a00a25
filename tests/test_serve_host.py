"""Host-side logic of the ensemble server (no GPU): workload generator,
footprint estimate, latency summary."""

import numpy as np

from magic_amd.serve import config5_requests, footprint_bytes, summarize, Result


def test_config5_workload_is_seeded_and_in_range():
    a = config5_requests()
    b = config5_requests()
    assert [(r.arrival, r.cfg) for r in a] == [(r.arrival, r.cfg) for r in b]
    assert len(a) == 64 and all(r.n_steps == 500 for r in a)
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) > 0)
    assert 20.0 < arr[-1] < 50.0          # 64 arrivals at 2 per second
    for r in a:
        c = r.cfg
        assert c["l_max"] in (32, 64, 85, 127) and c["n_r_max"] in (33, 49, 65)
        assert 1e-4 <= c["ek"] <= 1e-3
        assert 60.0 <= c["ra"] * c["ek"] <= 200.0


def test_footprint_grows_with_resolution():
    small = footprint_bytes(dict(l_max=32, n_r_max=33))
    big = footprint_bytes(dict(l_max=127, n_r_max=65))
    assert 0 < small < big < 4 << 30


def test_summary_percentiles():
    rs = [Result(i, {}, 3, 0.0, 0.0, 1.0, step_latencies=[0.01, 0.01, 0.02]) for i in range(4)]
    s = summarize(rs, 2.0)
    assert s["runs"] == 4 and s["steps"] == 12
    assert abs(s["p50_ms"] - 10.0) < 1e-9
    assert s["runs_per_hour"] == 7200.0
