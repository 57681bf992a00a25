"""The reference's module-level interface, driven the way its own callers
drive it (``configure`` -> ``init_fields.initialize_fields`` ->
``step_time.*`` -> ``output.*`` on ``fields.*``), against the reference's
fixtures; diagnostics on tensors passed in; the explicit-history views;
Fortran checkpoints and graph files against the files the reference wrote;
``main.run`` end to end."""

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


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


def test_module_level_facade_reproduces_reference_c1():
    """INTEGRATION.md's snippet: config 1 for 100 steps through the facade,
    energies every step (1e-9), fields (1e-10), and the step-100 e_kin / e_mag_oc
    tuples and get_dlm length scales of the reference (1e-10)."""
    import torch
    from magic_amd import config, fields, init_fields, output, step_time
    from magic_amd.dynamo import noise_factors
    case, g = _load("c1")
    config.configure(dict(case["cfg"]))
    init_fields.initialize_fields()
    run = config.default_run()
    coef, shape = noise_factors(run.p, run.res, case["noise"]["seed"], case["noise"]["amp"])
    # the reference's idiom: an in-place update of the module-level field
    fields.s_LMloc += torch.from_numpy(coef[:, None] * shape[None, :]).to(fields.s_LMloc.device)
    assert fields.s_LMloc.data_ptr() == run.f["s"].data_ptr()
    step_time.setup_initial_state()
    step_time.initialize_dt(case["dt"])
    dt = case["dt"]
    for n in range(1, case["steps"] + 1):
        dt = step_time.one_step(n, dt)
        ek = output.get_e_kin_full(fields.w_LMloc, fields.dw_LMloc, fields.z_LMloc)
        em = output.get_e_mag_oc_full(fields.b_LMloc, fields.db_LMloc, fields.aj_LMloc)
        e = np.array([ek.e_p, ek.e_t, em.e_p, em.e_t])
        assert np.all(np.abs(e - g["energies"][n - 1]) <= 1e-9 * np.abs(g["energies"][n - 1])), n
    for k in ("s", "z", "w", "p", "b", "aj"):
        got = getattr(fields, k + "_LMloc").cpu().numpy()
        assert _rel(got, g[f"{k}_step100"]) < 1e-10, k
    np.testing.assert_allclose(tuple(ek), g["ekin_full_step100"], rtol=1e-10)
    np.testing.assert_allclose(tuple(em), g["emag_full_step100"], rtol=1e-10, atol=1e-300)
    for sw, args in (("V", (fields.w_LMloc, fields.dw_LMloc, fields.z_LMloc)),
                     ("B", (fields.b_LMloc, fields.db_LMloc, fields.aj_LMloc))):
        dl, dlR, dm, dlc, pk, dlRc, pkR = output.get_dlm(*args, sw)
        np.testing.assert_allclose([dl, dm, dlc, pk], g[f"dlm{sw}_scalars_step100"], rtol=1e-10)
        np.testing.assert_allclose(np.stack([dlR.numpy(), dlRc.numpy(), pkR.numpy()]),
                                   g[f"dlm{sw}_radial_step100"], rtol=1e-10)
    config.configure({})


@pytest.mark.parametrize("name,step", [("c1", 100), ("dtchange", 30), ("minc2", 20),
                                       ("bprminc4", 20)])
def test_diagnostics_compute_on_the_tensors_given(name, step):
    """get_e_kin_full / get_e_mag_oc_full / get_dlm / spectra on arrays that are
    not the run's state (the reference's saved fields, host numpy): the
    reference's values (1e-12), and the oracle's spectra (1e-12); the run's own
    state is left alone."""
    import torch
    from magic_amd import config, output
    from oracle.dynamo import OracleRun
    case, g = _load(name)
    config.configure(dict(case["cfg"]))
    run = config.default_run()
    before = run.f["w"].clone()
    arr = {k: g[f"{k}_step{step}"] for k in ("w", "dw", "z", "b", "db", "aj")}
    ek = output.get_e_kin_full(arr["w"], arr["dw"], torch.from_numpy(arr["z"]))
    em = output.get_e_mag_oc_full(arr["b"], arr["db"], arr["aj"])
    np.testing.assert_allclose(tuple(ek), g[f"ekin_full_step{step}"], rtol=1e-12)
    np.testing.assert_allclose(tuple(em), g[f"emag_full_step{step}"], rtol=1e-12, atol=1e-300)
    assert output.get_e_kin(arr["w"], arr["dw"], None, arr["z"], None) == (ek.e_p, ek.e_t)
    for sw, args in (("V", (arr["w"], arr["dw"], arr["z"])), ("B", (arr["b"], arr["db"], arr["aj"]))):
        dl, dlR, dm, dlc, pk, dlRc, pkR = output.get_dlm(*args, sw)
        np.testing.assert_allclose([dl, dm, dlc, pk], g[f"dlm{sw}_scalars_step{step}"], rtol=1e-12)
        np.testing.assert_allclose(np.stack([dlR.numpy(), dlRc.numpy(), pkR.numpy()]),
                                   g[f"dlm{sw}_radial_step{step}"], rtol=1e-12)
    sp = output.get_spectra(**arr)
    orc = OracleRun(case["cfg"])
    for k, v in arr.items():
        getattr(orc, k)[:] = v
    osp = orc.spectra()
    for k in ("kin_l", "kin_m", "mag_l", "mag_m"):
        np.testing.assert_allclose(sp[k], osp[k], rtol=1e-12, atol=1e-14 * np.abs(osp[k]).max())
    assert torch.equal(run.f["w"], before)
    with pytest.raises(ValueError):
        output.get_e_kin_full(arr["w"][:, :5], arr["dw"], arr["z"])
    config.configure({})


def test_explicit_history_views_and_lcut():
    """dt_fields views alias the run's buffers in the reference's slot order
    (slot 1 = the history the next step reads; after a radial loop slot 0 is
    the new term), writes through them reach the run; the SHT facade's lcut
    changes nothing, as in the reference."""
    import torch
    from magic_amd import config, dt_fields, sht, step_time
    case, _ = _load("bench16")
    config.configure(dict(case["cfg"]))
    run = config.default_run()
    run.initialize_fields()
    step_time.setup_initial_state()
    step_time.initialize_dt(case["dt"])
    step_time.one_step(1, case["dt"])
    ex = dt_fields.dwdt.expl
    assert tuple(ex.shape) == tuple(run.f["w"].shape) + (2,)
    assert ex[:, :, 1].data_ptr() == run.expl_prev["w"].data_ptr()
    assert torch.equal(ex[..., 0], ex[..., 1])                 # after a whole step
    assert dt_fields.dsdt.old[..., 0].data_ptr() == run.old["s"].data_ptr()
    step_time.radial_loop()
    assert dt_fields.dwdt.expl[:, :, 0].data_ptr() == run.expl_cur["w"].data_ptr()
    assert dt_fields.dwdt.expl[:, :, 1].data_ptr() == run.expl_prev["w"].data_ptr()
    step_time.lm_loop()
    saved = run.expl_prev["z"].clone()
    dt_fields.dzdt.expl[:, :, 1] = 0.0
    assert not torch.count_nonzero(run.expl_prev["z"])
    dt_fields.dzdt.expl[:, :, 1] = saved
    assert torch.equal(run.expl_prev["z"], saved)
    x = run.f["s"][:, 3].clone()
    assert torch.equal(sht.scal_to_spat(x, lcut=4), sht.scal_to_spat(x))
    config.configure({})


def test_fortran_checkpoint_restart_matches_reference():
    """A run started from the reference's MagIC checkpoint the way main.run's
    fortran_restart does continues as the reference did (energies 1e-9 over 10
    steps, w 1e-10); this build's writer, read back, holds the same state."""
    import torch
    from magic_amd import checkpoint_io as cio
    from magic_amd.dynamo import DynamoRun
    g = np.load(os.path.join(GOLD, "files16.npz"))
    meta = json.load(open(os.path.join(GOLD, "files16.json")))
    run = DynamoRun(dict(meta["cfg"]))
    t, dt = cio.restart_from_fortran(run, os.path.join(GOLD, "files16_checkpoint.bin"))
    assert t == float(g["time"]) and dt == float(g["tscheme_dt"][0])
    for n in range(1, meta["cont"] + 1):
        dt = run.one_step(n, dt)
        e = np.array(run.energies())
        assert np.all(np.abs(e - g["energies_restart"][n - 1])
                      <= 1e-9 * np.abs(g["energies_restart"][n - 1])), n
    assert _rel(run.f["w"].cpu().numpy(), g["w_restart"]) < 1e-10
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "ck.bin")
        cio.write_checkpoint_fortran(path, run, run.time, 11)
        ck = cio.read_checkpoint(path)
        assert np.array_equal(ck.w, run.f["w"].cpu().numpy())
        assert np.array_equal(ck.derivatives["aj"]["expl"][0], run.expl_prev["j"].cpu().numpy())
    run.close()


def test_graph_file_matches_reference():
    """G_1 file of the state the reference wrote its graph file from: identical
    header and layout, every float32 value within 2e-6 of the largest value of
    its field (FP64 transforms rounded to float32 on both sides)."""
    import tempfile
    from magic_amd.dynamo import DynamoRun
    from magic_amd.graph_output import write_graph_file
    g = np.load(os.path.join(GOLD, "files16.npz"))
    meta = json.load(open(os.path.join(GOLD, "files16.json")))
    run = DynamoRun(dict(meta["cfg"]))
    ref = open(os.path.join(GOLD, "files16_G_1.bin"), "rb").read()
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
    scale = np.abs(oc).max(axis=(0, 2))          # per field of the outer core
    got_oc = slabs[0][:8 * N].reshape(N, 8, T * P)
    assert np.all(np.abs(got_oc - oc).max(axis=(0, 2)) <= 2e-6 * scale)
    ic_ref, ic_got = slabs[1][8 * N:], slabs[0][8 * N:]
    assert ic_ref.shape[0] == 3 * n_ic
    assert np.abs(ic_got - ic_ref).max() <= 2e-6 * np.abs(ic_ref).max()
    run.close()


def test_main_run_writes_the_reference_files(tmp_path):
    """main.run: fresh start, 20 steps with log and checkpoint cadence 10; its
    energies equal a one_step loop's bit for bit; the files are there with the
    reference's formats; a restart from the step-10 file ends at the same state."""
    import torch
    from magic_amd import checkpoint_io as cio
    from magic_amd import config, main
    from magic_amd.dynamo import DynamoRun
    cfg = dict(l_max=16, n_r_max=17, ek=1e-3, ra=1e5, pr=1.0, prmag=5.0, radratio=0.35,
               n_steps=20, dt=1e-4, log_every=10, checkpoint_every=10,
               output_dir=str(tmp_path), tag="t", n_graphs=1)
    config.configure(dict(cfg))
    e = main.run(dict(cfg))
    ref = DynamoRun(dict(cfg))
    ref.initialize_fields()
    ref.setup_initial_state()
    ref.initialize_dt(1e-4)
    dt = 1e-4
    for n in range(1, 21):
        dt = ref.one_step(n, dt)
    assert e == ref.energies()
    rows = (tmp_path / "log.csv").read_text().splitlines()
    assert rows[0].split(",")[:3] == ["step", "sim_time", "e_kin_pol"] and len(rows) == 4
    assert len((tmp_path / "e_kin.t").read_text().splitlines()) == 3
    assert len((tmp_path / "e_mag_oc.t").read_text().splitlines()) == 3
    assert (tmp_path / "checkpoint_000010.pt").exists() and (tmp_path / "checkpoint_000020.pt").exists()
    ck = cio.read_checkpoint(tmp_path / "checkpoint_end.t")
    assert ck.n_time_step == 21 and np.array_equal(ck.w, ref.f["w"].cpu().numpy())
    T, P, N = ref.p.n_theta, ref.p.n_phi, ref.p.n_r_max
    n_ic = (N + 1) // 2
    size = 4 + 64 + 40 + 44 + 4 * (T + N + n_ic) + 4 * T * P * (8 * N + 3 * n_ic)
    assert (tmp_path / "G_1.t").stat().st_size == size
    # restart from the step-10 file: same end state
    config.configure(dict(cfg))
    out2 = tmp_path / "restart"
    e2 = main.run(dict(cfg, restart=str(tmp_path / "checkpoint_000010.pt"),
                       output_dir=str(out2)))
    assert e2 == e
    assert torch.equal(config.default_run().f["w"], ref.f["w"])
    ref.close()
    config.configure({})


def test_movie_frames_match_reference():
    """Movie frames of the files16 state against the frames the reference's
    writer produced from the same state (its radial loop's grids): every float32
    value within 2e-6 of the frame's largest value (FP64 transforms on both
    sides, rounded to float32), for CMB and equatorial movies of Br, Bt, Vr, Vp
    and T."""
    from magic_amd.dynamo import DynamoRun
    from magic_amd.movie_output import movie_frame, parse_movie_string
    from test_movie_host import MOVIES, records
    g = np.load(os.path.join(GOLD, "files16.npz"))
    meta = json.load(open(os.path.join(GOLD, "files16.json")))
    run = DynamoRun(dict(meta["cfg"]))
    import torch
    for k, key in (("w", "w"), ("dw", "dw"), ("z", "z"), ("s", "s"), ("b", "b"), ("db", "db"),
                   ("aj", "aj")):
        run.f[key].copy_(torch.from_numpy(g[k]).to(run.device))
    for ms in MOVIES:
        ftype, nsurf, _, stem = parse_movie_string(ms, run.p)
        ref = records(open(os.path.join(GOLD, f"files16_{stem}.bin"), "rb").read())
        want = np.frombuffer(ref[11], ">f4").astype(np.float64)
        got = movie_frame(run, ftype, nsurf).astype(np.float64).reshape(-1)
        assert got.shape == want.shape, ms
        assert np.abs(got - want).max() <= 2e-6 * np.abs(want).max(), ms
        assert np.frombuffer(ref[10], ">f4")[0] == np.float32(g["time"])
    run.close()


def test_main_run_movie_cadence(tmp_path):
    """main.run with movies every 3 steps over 7 steps: two frames per file, and
    the frame of step 3 is the state before step 3 (after step 2), as the
    reference writes its last radial loop's grids."""
    from magic_amd import config, main
    from magic_amd.dynamo import DynamoRun
    from magic_amd.movie_output import movie_frame
    from test_movie_host import records
    cfg = dict(l_max=16, n_r_max=17, ek=1e-3, ra=1e5, pr=1.0, prmag=5.0, radratio=0.35,
               n_steps=7, dt=1e-4, log_every=7, output_dir=str(tmp_path), tag="t", n_graphs=0,
               movies=["Br CMB", "Vr EQ"], n_movie_step=3)
    config.configure(dict(cfg))
    main.run(dict(cfg))
    ref = DynamoRun(dict(cfg))
    ref.initialize_fields()
    ref.setup_initial_state()
    ref.initialize_dt(1e-4)
    dt = 1e-4
    for n in range(1, 3):
        dt = ref.one_step(n, dt)
    for stem, ft, ns in (("Br_CMB_mov", 1, 1), ("Vr_EQU_mov", 4, 2)):
        rec = records((tmp_path / f"{stem}.t").read_bytes())
        assert len(rec) == 10 + 2 * 2, stem
        want = movie_frame(ref, ft, ns).reshape(-1)
        assert np.array_equal(np.frombuffer(rec[11], ">f4"), want.astype(">f4")), stem
        assert np.frombuffer(rec[10], ">f4")[0] == np.float32(3e-4)
    ref.close()
    config.configure({})
