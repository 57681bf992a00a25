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


# the same with the finite-difference (double-curl) step, and with the anelastic step
EXTRA = {"fd": dict(radial_scheme="FD", fd_order=2, fd_order_bound=2),
         "anel": dict(strat=1.0, polind=2.0, g1=1.0)}


def _fresh(precision="fp64", extra=None):
    from magic_amd.dynamo import DynamoRun
    run = DynamoRun(dict(CASE["cfg"], precision=precision, **(extra or {})))
    run.initialize_dt(CASE["dt"])
    return run


def _started(precision="fp64", extra=None):
    run = _fresh(precision, extra)
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


@pytest.mark.parametrize("kind", ["fd", "anel"])
def test_restart_continues_bit_for_bit_fd_and_anelastic(tmp_path, kind):
    """A finite-difference (double-curl) run and an anelastic run restored from
    a .pt file written after step 12 continue bit for bit across the later dt
    changes, host-stepped and with the dt decision on the device."""
    import torch
    path = tmp_path / "restart.pt"
    extra = EXTRA[kind]
    ref = _started(extra=extra)
    dt = CASE["dt"]
    for n in range(1, CASE["at"] + 1):
        dt = ref.one_step(n, dt)
    ref.save_checkpoint(path)
    rebuilds_at = ref.n_rebuild
    for n in range(CASE["at"] + 1, CASE["steps"] + 1):
        dt = ref.one_step(n, dt)
    assert ref.n_rebuild > rebuilds_at, "the continuation should cross a dt change"
    a = _fresh(extra=extra)
    step, t = a.load_checkpoint(path)
    dt = a.dt_hist[0]
    for n in range(step + 1, CASE["steps"] + 1):
        dt = a.one_step(n, dt)
    assert a.dt_log == ref.dt_log[CASE["at"]:] and a.time == ref.time
    for k in ("s", "z", "w", "dw", "ddw", "p", "b", "aj"):
        assert torch.equal(a.f[k], ref.f[k]), k
    b = _fresh(extra=extra)
    b.use_graphs = True
    b.load_checkpoint(path)
    b.run_steps(CASE["steps"] - CASE["at"], b.dt_hist[0])
    assert b.dt_log == ref.dt_log[CASE["at"]:]
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


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _gold(name):
    import json
    import os
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    return (np.load(os.path.join(gold, f"{name}.npz")),
            json.load(open(os.path.join(gold, f"{name}.json"))), os.path.join(gold, name))


@pytest.mark.parametrize("name", ["filesrot8", "filesbprrot8"])
@pytest.mark.parametrize("kind", ["fortran", "pt"])
def test_inner_core_restart_from_reference_files_matches_reference(name, kind):
    """Conducting, rotating inner core (CNAB2 and BPR353): a run restarted from
    the reference's own Fortran checkpoint (main.run's fortran_restart sequence)
    or its .pt restart file, both written after 10 steps, continues as the
    reference did after the same restart: energies within 1e-9 and omega_ic
    within 1e-10 every step, w and b_ic within 1e-10 after 10 steps."""
    from magic_amd import checkpoint_io as cio
    from magic_amd.dynamo import DynamoRun
    g, meta, stem = _gold(name)
    run = DynamoRun(dict(meta["cfg"]))
    if kind == "fortran":
        t, dt = cio.restart_from_fortran(run, stem + "_checkpoint.bin")
        start = 0
    else:
        start, t = run.load_checkpoint(stem + "_restart.pt")
        dt = run.dt_hist[0]
    assert t == float(g["time"])
    for i, n in enumerate(range(start + 1, start + meta["cont"] + 1)):
        dt = run.one_step(n, dt)
        e, ge = np.array(run.energies()), g[f"energies_{kind}"][i]
        assert np.all(np.abs(e - ge) <= 1e-9 * np.abs(ge)), (n, e, ge)
        assert run.omega_ic == pytest.approx(g[f"omega_{kind}"][i], rel=1e-10), n
    assert _rel(run.f["w"].cpu().numpy(), g[f"w_{kind}"]) < 1e-10
    assert _rel(run.fic["b_ic"].cpu().numpy(), g[f"b_ic_{kind}"]) < 1e-10
    run.close()


@pytest.mark.parametrize("scheme", ["CNAB2", "BPR353"])
def test_inner_core_restart_continues_bit_for_bit(tmp_path, scheme):
    """A rotating conducting inner core saved after 6 steps and restored into a
    fresh run continues bit for bit (fields, inner fields, rotation state) like
    the run that was never interrupted; the Fortran checkpoint it writes reads
    back to the same fields and (CNAB2) histories."""
    import torch
    from magic_amd import checkpoint_io as cio
    from magic_amd.dynamo import DynamoRun
    cfg = dict(l_max=8, n_r_max=17, n_cheb_max=15, n_cheb_ic_max=7, sigma_ratio=1.0,
               kbotb=3, nRotIC=1, dtmax=2e-4, time_scheme=scheme)
    ref = DynamoRun(cfg)
    ref.initialize_fields(noise_amp=1e-2, noise_seed=5)
    ref.setup_initial_state()
    ref.initialize_dt(2e-4)
    dt = 2e-4
    for n in range(1, 7):
        dt = ref.one_step(n, dt)
    path = tmp_path / "r.pt"
    ref.save_checkpoint(path)
    ckf = tmp_path / "ck.bin"
    cio.write_checkpoint_fortran(ckf, ref, ref.time, 7)
    ck = cio.read_checkpoint(ckf)
    assert ck.family == ("DIRK" if ref.dirk else "MULTISTEP")
    assert np.array_equal(ck.b_ic, ref.fic["b_ic"].cpu().numpy())
    if not ref.dirk:
        assert np.array_equal(ck.derivatives["aj_ic"]["expl"][0],
                              ref.expl_ic[1 - ref._par]["j"].cpu().numpy())
        assert ck.domega_ic_dt["expl"][0] == float(ref.rot[7 - ref._par])
    for n in range(7, 13):
        dt = ref.one_step(n, dt)
    a = DynamoRun(cfg)
    step, _ = a.load_checkpoint(path)
    dt = a.dt_hist[0]
    for n in range(step + 1, 13):
        dt = a.one_step(n, dt)
    assert a.time == ref.time and a.dt_log == ref.dt_log[6:]
    for k in ("s", "z", "w", "p", "b", "aj"):
        assert torch.equal(a.f[k], ref.f[k]), k
    for k in ("b_ic", "aj_ic"):
        assert torch.equal(a.fic[k], ref.fic[k]), k
    assert torch.equal(a.rot[:10], ref.rot[:10])
    for r in (ref, a):
        r.close()


def test_inner_core_graph_file_matches_reference():
    """G_1 file of a conducting, rotating inner-core state: the reference's
    header (l_cond_ic flag, sigma_ratio) byte for byte, every float32 value
    within 2e-6 of the largest value of its field."""
    import tempfile
    import os
    from magic_amd.dynamo import DynamoRun
    from magic_amd.graph_output import write_graph_file
    g, meta, stem = _gold("filesrot8")
    run = DynamoRun(dict(meta["cfg"]))
    ref = open(stem + "_G_1.bin", "rb").read()
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "G_1.t")
        write_graph_file(path, float(g["time"]), run,
                         {k: g[k] for k in ("w", "dw", "z", "s", "p", "b", "db", "aj")})
        got = open(path, "rb").read()
    assert len(got) == len(ref)
    hdr = 4 + 64 + 4 * 10 + 4 * 11
    assert got[:hdr] == ref[:hdr]
    a = np.frombuffer(got[hdr:], ">f4").astype(np.float64)
    b = np.frombuffer(ref[hdr:], ">f4").astype(np.float64)
    T, P, N = run.p.n_theta, run.p.n_phi, run.p.n_r_max
    n_ic = (N + 1) // 2
    axes = T + N + n_ic
    np.testing.assert_allclose(a[:axes], b[:axes], rtol=2e-7, atol=0)
    slabs = (a[axes:].reshape(-1, T * P), b[axes:].reshape(-1, T * P))
    oc = slabs[1][:8 * N].reshape(N, 8, T * P)
    scale = np.abs(oc).max(axis=(0, 2))
    got_oc = slabs[0][:8 * N].reshape(N, 8, T * P)
    assert np.all(np.abs(got_oc - oc).max(axis=(0, 2)) <= 2e-6 * scale)
    ic_ref, ic_got = slabs[1][8 * N:], slabs[0][8 * N:]
    assert np.abs(ic_got - ic_ref).max() <= 2e-6 * np.abs(ic_ref).max()
    run.close()
