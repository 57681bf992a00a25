"""Host-side logic of the ensemble server (no GPU): workload generator,
footprint estimate, latency summary, row-job packing."""

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


def test_row_job_packing_matches_struct_layout():
    """The numpy packing of row-operator jobs equals MhRowJob records built one
    by one with ctypes, and the tile list enumerates (job, 32-row chunk,
    64-column chunk) in job order."""
    import ctypes as C

    from magic_amd import _lib
    from magic_amd.dynamo import pack_row_jobs
    ls = np.arange(40)
    specs = [dict(X=0x1000, ldx=33, rows=0x2000 + 4 * ls * (ls + 1) // 2, nrows=ls + 1,
                  MT=0x3000 + 8 * ls * 777, K=33, R=99, Rpad=128, Y=[0x10, 0x20, 0x30],
                  ldy=33, rseg=33),
             dict(X=0x5000, ldx=33, rows=0, nrows=561, MT=0x6000, K=33, R=33, Rpad=64,
                  Y=[0x40], ldy=33, rseg=33)]
    arr, tiles = pack_row_jobs(specs, n_xcd=1)
    ref, ref_tiles = [], []
    for sp in specs:
        for i in range(np.size(sp["nrows"])):
            j = _lib.RowJob()
            pick = lambda v: int(np.asarray(v).reshape(-1)[i]) if np.size(v) > 1 else int(v)
            j.X, j.rows, j.MT = sp["X"], pick(sp["rows"]) or None, pick(sp["MT"])
            for k, y in enumerate(sp["Y"]):
                j.Y[k] = y
            j.ldx, j.ldy, j.K, j.R, j.Rpad, j.rseg = (sp[k] for k in ("ldx", "ldy", "K", "R", "Rpad", "rseg"))
            j.nrows = pick(sp["nrows"])
            ref.append(j)
    raw = bytes((_lib.RowJob * len(ref))(*ref))
    assert arr.tobytes() == raw
    for ji, j in enumerate(ref):
        for r0 in range(0, j.nrows, 32):
            for c0 in range(0, j.R, 64):
                ref_tiles.append((ji, r0, c0))
    assert tiles.tolist() == [list(t) for t in ref_tiles]
    assert C.sizeof(_lib.RowJob) == arr.dtype.itemsize
    # the default order is a permutation of the same tiles
    arr8, tiles8 = pack_row_jobs(specs)
    assert arr8.tobytes() == raw
    assert sorted(map(tuple, tiles8.tolist())) == sorted(ref_tiles)


def test_xcd_tile_order_keeps_each_job_on_one_xcd():
    """Workgroup i runs on XCD i mod 8: every tile of an unsplit job gets the
    same i mod 8, a job larger than 1/64 of the tiles is cut at row-chunk
    boundaries, and the eight XCDs get equal tile counts up to the largest
    unit."""
    from magic_amd.dynamo import xcd_tile_order
    ls = np.arange(384)
    # per-degree solves at l_max=383 (4 fields, 3 column tiles per 32-row chunk)
    per = np.repeat(3 * (-(-(ls + 1) // 32)), 4)
    nc = np.full(per.size, 3)
    o = xcd_tile_order(per, nc)
    assert sorted(o.tolist()) == list(range(int(per.sum())))
    job = np.repeat(np.arange(per.size), per)[o]
    pos = np.arange(o.size)
    for j in range(per.size):
        assert np.unique(pos[job == j] % 8).size == 1
    counts = np.bincount(pos % 8, minlength=8)
    assert counts.max() - counts.min() <= 1
    # two derivative jobs over all 73 920 rows: cut into pieces of 72 tiles
    per = np.array([2310, 2310])
    o = xcd_tile_order(per, np.array([1, 1]))
    assert sorted(o.tolist()) == list(range(4620))
    first = o[np.arange(0, o.size, 8)]          # XCD 0's tiles in dispatch order
    runs = np.split(first, np.nonzero(np.diff(first) != 1)[0] + 1)
    assert all(r.size <= 72 for r in runs) and sum(r.size for r in runs) == first.size
