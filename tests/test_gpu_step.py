"""GPU parity of the full CNAB2 dynamo step against the reference trajectory.

The golden fixtures (``tests/golden/traj_*.npz``) were written by running the
reference's own ``step_time.one_step`` on the CPU (``make_golden.py``); the
oracle reproduces them to ~1e-13.  The bar (north star): spectral
coefficients within 1e-10 relative and energies within 1e-9 relative.
"""

import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    with open(os.path.join(GOLD, f"traj_{name}.json")) as f:
        case = json.load(f)
    return case, np.load(os.path.join(GOLD, f"traj_{name}.npz"))


def _run(case):
    from magic_amd.dynamo import DynamoRun
    run = DynamoRun(case["cfg"])
    nz = case["noise"]
    if nz:
        run.initialize_fields(noise_amp=nz["amp"], noise_seed=nz["seed"])
    else:
        run.initialize_fields()
    run.setup_initial_state()
    run.initialize_dt(case["dt"])
    return run


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("name", ["bench16", "c1"])
def test_step_matches_reference_trajectory(name):
    case, g = _load(name)
    run = _run(case)
    for n in range(1, case["steps"] + 1):
        dt = run.one_step(n, case["dt"])
        assert dt == pytest.approx(g["dt"][n - 1], rel=1e-12)
        e = np.array(run.energies())
        ge = g["energies"][n - 1]
        assert np.all(np.abs(e - ge) <= 1e-9 * np.abs(ge)), (n, e, ge)
        if n in case["field_steps"]:
            for k in ("s", "z", "w", "p", "b", "aj"):
                got = run.f[k].cpu().numpy()
                assert _rel(got, g[f"{k}_step{n}"]) < 1e-10, (n, k, _rel(got, g[f"{k}_step{n}"]))


def test_first_step_explicit_terms():
    """Explicit terms after the first radial loop (get_nl + get_td in context)."""
    case, g = _load("bench16")
    run = _run(case)
    run.one_step(1, case["dt"])
    for k in ("s", "z", "w", "p", "b", "j"):
        ref = g[f"expl_{k}_step1"]
        got = run.expl_prev[k].cpu().numpy()
        scale = max(np.abs(ref).max(), 1e-300)
        assert np.abs(got - ref).max() / scale < 1e-10 or np.abs(ref).max() == 0.0, k
