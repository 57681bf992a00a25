"""Restart files in the reference's .pt layout (``main.py:189-267``): a run
restored from a file written after step k continues bit for bit like the run
that was never interrupted, host-stepped and with the dt decision on the device."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# l_max=16, n_r=33 with strong noise: the Courant rule changes dt (and rebuilds the
# operators) several times in 30 steps, so the restart crosses dt changes
CASE = dict(cfg=dict(l_max=16, n_r_max=33, ek=1e-3, ra=1e5, pr=1.0, prmag=5.0,
                     radratio=0.35, dtmax=1e-3),
            dt=1e-3, noise=dict(amp=1.0, seed=7), steps=30, at=12)


def _fresh(precision="fp64"):
    from magic_amd.dynamo import DynamoRun
    run = DynamoRun(dict(CASE["cfg"], precision=precision))
    run.initialize_dt(CASE["dt"])
    return run


def _started(precision="fp64"):
    run = _fresh(precision)
    run.initialize_fields(noise_amp=CASE["noise"]["amp"], noise_seed=CASE["noise"]["seed"])
    run.setup_initial_state()
    return run


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_restart_continues_bit_for_bit(tmp_path, precision):
    import torch
    path = tmp_path / "restart.pt"
    ref = _started(precision)
    dt = CASE["dt"]
    for n in range(1, CASE["at"] + 1):
        dt = ref.one_step(n, dt)
    ref.save_checkpoint(path)
    t_at, dt_at, rebuilds_at = ref.time, dt, ref.n_rebuild
    for n in range(CASE["at"] + 1, CASE["steps"] + 1):
        dt = ref.one_step(n, dt)
    assert ref.n_rebuild > rebuilds_at, "the continuation should cross a dt change"

    # the file: the reference's keys, CNAB2 slot counts, loadable without unpickling code
    st = torch.load(path, weights_only=True)
    assert st["step"] == CASE["at"] and st["sim_time"] == t_at
    assert float(st["tscheme_dt"][0]) == dt_at
    L, N = ref.f["s"].shape
    assert st["fields.w_LMloc"].shape == (L, N) and st["dt.dwdt.expl"].shape == (L, N, 2)
    assert st["dt.dsdt.old"].shape == (L, N, 1) and st["dt.dbdt.impl"].shape == (L, N, 1)
    assert torch.equal(st["dt.dzdt.expl"][..., 0], st["dt.dzdt.expl"][..., 1])
    assert not torch.count_nonzero(st["fields.xi_LMloc"])

    # host-stepped continuation from the file
    a = _fresh(precision)
    step, t = a.load_checkpoint(path)
    assert (step, t) == (CASE["at"], t_at)
    dt = a.dt_hist[0]
    for n in range(step + 1, CASE["steps"] + 1):
        dt = a.one_step(n, dt)
    assert a.dt_log == ref.dt_log[CASE["at"]:]
    assert a.time == ref.time
    for k in ("s", "ds", "z", "dz", "w", "dw", "ddw", "p", "dp", "b", "db", "ddb", "aj", "dj"):
        assert torch.equal(a.f[k], ref.f[k]), k
    assert a.energies() == ref.energies()
    # the time arrays too (the reference's restart test compares them slot by slot)
    for k in ("s", "z", "w", "p", "b", "j"):
        assert torch.equal(a.old[k], ref.old[k]), ("old", k)
        assert torch.equal(a.impl[k], ref.impl[k]), ("impl", k)
        assert torch.equal(a.expl_prev[k], ref.expl_prev[k]), ("expl", k)

    # device-controlled continuation (one HIP graph per step)
    b = _fresh(precision)
    b.use_graphs = True
    b.load_checkpoint(path)
    b.run_steps(CASE["steps"] - CASE["at"], b.dt_hist[0])
    assert b.dt_log == ref.dt_log[CASE["at"]:] and b.time == ref.time
    for k in ("s", "z", "w", "p", "b", "aj"):
        assert torch.equal(b.f[k], ref.f[k]), k
    for r in (ref, a, b):
        r.close()


def test_restart_refuses_queued_device_steps(tmp_path):
    run = _started()
    run.ctl_begin(CASE["dt"])
    run.step_async()
    with pytest.raises(RuntimeError, match="ctl_end"):
        run.save_checkpoint(tmp_path / "x.pt")
    run.ctl_end()
    run.save_checkpoint(tmp_path / "x.pt")
    run.close()


def test_load_after_device_steps_resets_step_control(tmp_path):
    """A run that stepped past the file with device step control, then loads
    the earlier file, saves and continues from the restored state (not from the
    pre-load device state): same steps as a run restored in a fresh object."""
    import torch
    path, path2 = tmp_path / "at12.pt", tmp_path / "again.pt"
    src = _started()
    src.run_steps(CASE["at"], CASE["dt"])
    src.save_checkpoint(path)
    src.use_graphs = True
    src.run_steps(8, src.dt_hist[0])           # past the file
    step, _ = src.load_checkpoint(path)
    src.save_checkpoint(path2)                 # no stale queued steps
    assert torch.load(path2, weights_only=True)["step"] == step == CASE["at"]
    src.run_steps(CASE["steps"] - step, src.dt_hist[0])
    ref = _fresh()
    ref.load_checkpoint(path)
    ref.run_steps(CASE["steps"] - step, ref.dt_hist[0])
    assert src.n_step == ref.n_step == CASE["steps"] and src.time == ref.time
    for k in ("s", "z", "w", "p", "b", "aj"):
        assert torch.equal(src.f[k], ref.f[k]), k
    for r in (src, ref):
        r.close()
