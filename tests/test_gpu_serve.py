"""Ensemble server on the GPU: continuous admission, overlapped runs, and the
same result for every run alone and in the mix."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests():
    from magic_amd.serve import Request
    base = dict(pr=1.0, prmag=5.0, radratio=0.35, dtmax=1e-4)
    cfgs = [dict(base, l_max=16, n_r_max=17, ek=1e-3, ra=1e5),
            dict(base, l_max=32, n_r_max=33, ek=5e-4, ra=2e5),
            dict(base, l_max=16, n_r_max=25, ek=2e-4, ra=5e5),
            dict(base, l_max=32, n_r_max=17, ek=1e-3, ra=8e4)]
    # staggered arrivals: later runs are admitted while earlier ones step
    return [Request(i, c, n_steps=12 + 3 * i, arrival=0.05 * i, noise_amp=1e-5, noise_seed=7 + i)
            for i, c in enumerate(cfgs)]


def test_each_run_is_the_same_alone_and_in_the_mix():
    import torch
    from magic_amd.serve import EnsembleServer, serve_sequential
    mix, summary = EnsembleServer(max_slots=3).serve(_requests(), keep_fields=True)
    alone, _ = serve_sequential(_requests(), keep_fields=True)
    assert summary["runs"] == 4 and summary["steps"] == sum(12 + 3 * i for i in range(4))
    by_id = {r.rid: r for r in alone}
    for r in mix:
        a = by_id[r.rid]
        assert r.energies == a.energies, r.rid
        for k, v in r.final_fields.items():
            assert torch.equal(v, a.final_fields[k]), (r.rid, k)
    # the mix overlapped: with three slots some round held more than one run
    assert len(mix) == 4


def test_admission_respects_slots_and_budget():
    from magic_amd.serve import EnsembleServer, footprint_bytes
    reqs = _requests()
    budget = max(footprint_bytes(r.cfg) for r in reqs) + 1
    srv = EnsembleServer(max_slots=4, hbm_budget=budget)
    done, summary = srv.serve(reqs)
    assert summary["runs"] == 4
    assert max(n for _, n in srv.occupancy) == 1  # only one run fits the budget at a time


def test_runs_of_one_resolution_overlapping_match_alone():
    """Three runs of the same resolution (l_max=32, n_r=33; parameters from the
    ensemble workload) admitted together, so their kernels share the compute
    units: each must end bit-identical to the same run alone."""
    import torch
    from magic_amd.serve import EnsembleServer, Request, serve_sequential
    base = dict(l_max=32, n_r_max=33, pr=1.0, prmag=5.0, radratio=0.35, dtmax=1e-4)
    cfgs = [dict(base, ek=1.92e-4, ra=8.43e5), dict(base, ek=3.01e-4, ra=5.37e5),
            dict(base, ek=1.82e-4, ra=8.58e5)]

    def reqs():
        return [Request(i, c, n_steps=60, arrival=0.0, noise_amp=1e-5, noise_seed=1006 + i)
                for i, c in enumerate(cfgs)]

    srv = EnsembleServer(max_slots=3)
    mix, _ = srv.serve(reqs(), realtime=False, keep_fields=True)
    alone, _ = serve_sequential(reqs(), keep_fields=True)
    assert max(n for _, n in srv.occupancy) == 3
    by_id = {r.rid: r for r in alone}
    for r in mix:
        a = by_id[r.rid]
        assert r.energies == a.energies, r.rid
        for k, v in r.final_fields.items():
            assert torch.equal(v, a.final_fields[k]), (r.rid, k)


def test_config5_resolutions_alone_and_in_the_mix():
    """Requests of the ensemble workload itself (config5_requests) covering
    l_max 64, 85 and 127 and n_r 49 and 65, 30 steps each, served with
    overlapping admission: each run ends bit-identical to the same run alone."""
    import torch
    from magic_amd.serve import EnsembleServer, Request, config5_requests, serve_sequential
    pool = config5_requests()
    picked, seen = [], set()
    for r in pool:
        key = (r.cfg["l_max"], r.cfg["n_r_max"])
        if r.cfg["l_max"] in (64, 85, 127) and r.cfg["n_r_max"] in (49, 65) and key not in seen:
            seen.add(key)
            picked.append(r)
    assert {k[0] for k in seen} == {64, 85, 127} and {k[1] for k in seen} == {49, 65}
    picked = picked[:6]

    def reqs():
        return [Request(i, dict(r.cfg), n_steps=30, arrival=0.02 * i, noise_amp=1e-5,
                        noise_seed=r.noise_seed) for i, r in enumerate(picked)]

    srv = EnsembleServer(max_slots=4)
    mix, summary = srv.serve(reqs(), keep_fields=True)
    alone, _ = serve_sequential(reqs(), keep_fields=True)
    assert summary["runs"] == len(picked)
    assert max(n for _, n in srv.occupancy) >= 2
    by_id = {r.rid: r for r in alone}
    for r in mix:
        a = by_id[r.rid]
        assert r.energies == a.energies, (r.rid, r.cfg)
        for k, v in r.final_fields.items():
            assert torch.equal(v, a.final_fields[k]), (r.rid, k)
