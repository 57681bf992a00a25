"""Host-side bounds check of the recorded launch plans (no GPU).

A DynamoRun is built on a CPU stand-in device with every C-ABI call recorded
instead of run; the test then checks, from the recorded arguments, that every
element range the rebuild kernels (matrix build, Gauss-Jordan inverse, fold
GEMM, fill), the row-operator jobs and the step control touch lies inside a
tensor of the run or its resolution, for every resolution the ensemble
workload serves and the headline one.
"""

import ctypes as C

import numpy as np
import pytest
import torch

from magic_amd import _lib
from magic_amd import device as devmod
from magic_amd import dynamo as dyn


class _NoStream:
    def synchronize(self):
        pass


@pytest.fixture
def cpu_standin(monkeypatch):
    def require(d):
        _lib.load()
        return torch.device("cpu")
    monkeypatch.setattr(devmod, "require_gpu", require)
    monkeypatch.setattr(dyn, "require_gpu", require)
    monkeypatch.setattr(devmod, "stream_handle", lambda: C.c_void_p(0))
    monkeypatch.setattr(dyn, "stream_handle", lambda: C.c_void_p(0))
    monkeypatch.setattr(torch.cuda, "current_stream", lambda *a, **k: _NoStream())
    monkeypatch.setattr(torch.Tensor, "pin_memory", lambda self: self)
    monkeypatch.setenv("MAGIC_AMD_TUNE", "0")  # no launch-shape timing without a GPU


def _extents(run):
    out = []

    def add(v):
        if isinstance(v, torch.Tensor):
            st = v.untyped_storage()
            out.append((st.data_ptr(), st.data_ptr() + st.nbytes()))
        elif isinstance(v, dict):
            for w in v.values():
                add(w)
        elif isinstance(v, (tuple, list)):
            for w in v:
                add(w)

    for obj in (run, run.res):
        for v in vars(obj).values():
            add(v)
    return out


def _inside(ext, lo, hi, what):
    assert any(a <= lo and hi <= b for a, b in ext), f"{what}: [{lo:#x}, {hi:#x}) outside every tensor"


def _check_rebuild_calls(calls, ext):
    n = 0
    for _, name, args in calls:
        if name == "mh_build_matrices":
            a = C.cast(args[0], C.POINTER(_lib.BuildArgs)).contents
            N, lp1 = a.N, a.l_max + 1
            if a.kind == 6:  # all systems: stacked N x N systems, (w,p), l=0 pressure
                _inside(ext, a.out, a.out + 8 * a.nsys * lp1 * N * N, "build stack")
                _inside(ext, a.out_wp, a.out_wp + 8 * lp1 * 4 * N * N, "build (w,p)")
                _inside(ext, a.out_p0, a.out_p0 + 8 * N * N, "build p0")
            else:
                size = {4: lp1 * 4 * N * N, 5: N * N}.get(a.kind, lp1 * N * N)
                _inside(ext, a.out, a.out + 8 * size, f"build kind {a.kind}")
            for p in (a.R, a.D1, a.D2, a.D3):
                _inside(ext, p, p + 8 * N * N, "Chebyshev matrix")
            for p in (a.or1, a.or2):
                _inside(ext, p, p + 8 * N, "radial factor")
        elif name == "mh_gj_inverse":
            A, out, work, scal, n_, nb, pre, info = args[:8]
            assert n_ <= 208  # register kernels (no work space)
            _inside(ext, A, A + 8 * nb * n_ * n_, "inverse input")
            _inside(ext, out, out + 8 * nb * n_ * n_, "inverse output")
            _inside(ext, info, info + 4 * nb, "pivot flags")
        elif name == "mh_gemm_f64":
            g = C.cast(args[0], C.POINTER(_lib.GemmArgs)).contents
            last = g.batch - 1
            for p, sb, sr, sc, rows, cols in ((g.A, g.sAb, g.sAr, g.sAc, g.M, g.K),
                                              (g.B, g.sBb, g.sBr, g.sBc, g.K, g.Nn),
                                              (g.C, g.sCb, g.sCr, g.sCc, g.M, g.Nn)):
                hi = p + 8 * (last * sb + (rows - 1) * sr + (cols - 1) * sc + 1)
                _inside(ext, p, hi, "gemm operand")
        elif name == "mh_fill_f64":
            p, nf, count, stride = args[:4]
            for k in range(count):
                _inside(ext, p + 8 * k * stride, p + 8 * (k * stride + nf), f"fill block {k}")
        else:
            continue
        n += 1
    return n


def _check_row_jobs(run, ext):
    for key in ("jobs1", "jobs2", "jobs_td", "jobs_setup", "jobs_setup2"):
        if not hasattr(run, key):
            continue
        dj, tile_lists = getattr(run, key)
        jobs = np.frombuffer(dj.numpy().tobytes(), dtype=dyn._ROWJOB_DTYPE)
        for tc, (tl, nt) in tile_lists.items():
            tiles = tl.numpy().reshape(-1, 3)
            assert len(tiles) == nt and tiles[:, 0].max() < len(jobs)
            # every tile starts inside its job: rows below nrows, columns below R
            assert np.all(tiles[:, 1] < jobs["nrows"][tiles[:, 0]])
            assert np.all(tiles[:, 2] < jobs["R"][tiles[:, 0]]) and np.all(tiles[:, 2] % tc == 0)
        for j in jobs:
            nrows, ldx, ldy, K, R, Rpad, rseg = (int(j[k]) for k in
                                                 ("nrows", "ldx", "ldy", "K", "R", "Rpad", "rseg"))
            if j["rows"]:
                _inside(ext, int(j["rows"]), int(j["rows"]) + 4 * nrows, f"{key} rows")
                last = int(np.frombuffer(C.string_at(int(j["rows"]), 4 * nrows), np.int32).max())
            else:
                last = nrows - 1
            _inside(ext, int(j["X"]), int(j["X"]) + 8 * (2 * last * ldx + 2 * K), f"{key} X")
            _inside(ext, int(j["MT"]), int(j["MT"]) + 8 * (-(-K // 4) * 4) * Rpad, f"{key} MT")
            for s in range(-(-R // rseg)):
                y = int(j["Y"][s])
                _inside(ext, y, y + 8 * (2 * last * ldy + 2 * rseg), f"{key} Y{s}")


@pytest.mark.parametrize("l_max,n_r", [(16, 17), (32, 33), (64, 49), (85, 65), (127, 65), (383, 97)])
def test_recorded_launches_stay_inside_their_tensors(cpu_standin, l_max, n_r):
    cfg = dict(l_max=l_max, n_r_max=n_r, ek=1e-4, ra=1e6, pr=1.0, prmag=5.0, radratio=0.35,
               dtmax=1e-4)
    with _lib.record():
        run = dyn.DynamoRun(cfg, device="cpu")
        run.initialize_dt(cfg["dtmax"])
        run._ctl_push(cfg["dtmax"])
    ext = _extents(run)
    _check_row_jobs(run, ext)
    base = devmod.ptr(run._ctl["state"])
    with _lib.record() as rec:
        run._rebuild_launches(0.3)
        run._rebuild_launches(0.0, (base + _lib.StepState.wimp_lin0.offset,
                                    base + _lib.StepState.rebuild.offset))
        run._ctl_launch()
    assert _check_rebuild_calls(rec.calls, ext) == 2 * 10  # ten launches per rebuild
    a = C.cast(rec.calls[-1][2][0], C.POINTER(_lib.StepCtlArgs)).contents
    _inside(ext, a.cfl, a.cfl + 16 * a.N, "Courant maxima")
    _inside(ext, a.info, a.info + 4 * a.info_rows * a.info_cols, "pivot flags")
    _inside(ext, a.dt_log, a.dt_log + 8 * a.log_cap, "dt log")
    _inside(ext, a.state, a.state + C.sizeof(_lib.StepState), "step state")


def test_run_is_freed_with_its_last_reference(cpu_standin):
    """Recording launch plans (sections, timers) leaves no reference cycle: a run
    and its buffers (and, on the GPU, its captured graphs) are released when the
    last reference goes, never later at an arbitrary garbage collection, which
    could fall inside another run's graph capture."""
    import gc
    import weakref
    cfg = dict(l_max=16, n_r_max=17, ek=1e-3, ra=1e5, pr=1.0, prmag=5.0, radratio=0.35,
               dtmax=1e-4)
    was = gc.isenabled()
    gc.disable()
    try:
        with _lib.record():
            run = dyn.DynamoRun(cfg, device="cpu")
            run.initialize_dt(cfg["dtmax"])
            run._ctl_push(cfg["dtmax"])
        for kind in ("step", "radial", "lm"):
            run._plan(kind)
        ref = weakref.ref(run)
        del run
        assert ref() is None
    finally:
        if was:
            gc.enable()


def test_mixed_precision_tables_are_converted_once_outside_the_plans(cpu_standin):
    """The FP32 copies of the Legendre tables are made when a mixed-precision
    run is built (four conversions), never inside a recorded step plan, where
    they would run again with every replay while other runs read the tables."""
    cfg = dict(l_max=16, n_r_max=17, ek=1e-3, ra=1e5, pr=1.0, prmag=5.0, radratio=0.35,
               dtmax=1e-4, precision="mixed")
    with _lib.record() as built:
        run = dyn.DynamoRun(cfg, device="cpu")
        run.initialize_dt(cfg["dtmax"])
        run._ctl_push(cfg["dtmax"])
    assert sum(name == "mh_tables_to_f32" for _, name, _ in built.calls) == 4
    for kind in ("radial", "lm", "step"):
        for par in (0, 1):
            run._par = par
            _, plan = run._plan(kind)
            assert not any(name == "mh_tables_to_f32" for _, name, _ in plan), (kind, par)
            if kind != "lm":
                assert any(name == "mh_leg_synth" for _, name, _ in plan)
    # a second run of the resolution reuses the same tables
    with _lib.record() as again:
        dyn.DynamoRun(cfg, device="cpu", res=run.res)
    assert not any(name == "mh_tables_to_f32" for _, name, _ in again.calls)


@pytest.mark.parametrize("scheme", ["CNAB2", "BPR353"])
def test_conducting_inner_core_launches_stay_inside_their_tensors(cpu_standin, scheme):
    """Conducting inner core: the coupled B/J row-operator jobs (N + n_ic long
    right-hand sides, outer and inner outputs), the inner derivative jobs of the
    set-up, and the arguments of mh_ic_rhs / mh_ic_impl in every recorded lm
    plan (each BPR353 stage) lie inside the run's tensors."""
    cfg = dict(l_max=16, n_r_max=17, n_cheb_ic_max=7, sigma_ratio=1.0, kbotb=3, ek=1e-3,
               ra=1e5, pr=1.0, prmag=5.0, radratio=0.35, dtmax=1e-4, time_scheme=scheme)
    with _lib.record() as rec:
        run = dyn.DynamoRun(cfg, device="cpu")
        run.initialize_dt(cfg["dtmax"])
        run._rebuild_launches(0.5e-4)
    assert run.cond_ic
    ext = _extents(run)
    _check_row_jobs(run, ext)
    # the device rebuild of the coupled operators: affine assembly, inverses, folds
    aff = [a for _, name, a in rec.calls if name == "mh_affine_f64"]
    for A0, A1, out, cnt in (a[:4] for a in aff):
        for v in (A0, A1, out):
            _inside(ext, v, v + 8 * cnt, "affine operand")
    assert _check_rebuild_calls(rec.calls, ext) == 10 + 1 + 4 + 4  # + inverse, 4 folds, 4 fills
    L, N, nic = run.L, run.N, run.p.n_r_ic_max
    stages = (1, 2, 3, 4) if run.dirk else (0,)
    n_ic_calls = 0
    for k in stages:
        run.stage = k
        _, plan = run._plan("lm")
        for _, name, args in plan:
            if name == "mh_ic_rhs":
                a = C.cast(args[0], C.POINTER(_lib.IcRhsArgs)).contents
                assert (a.L, a.N, a.n_ic) == (L, N, nic)
                for sy in a.sys:
                    for f in ("old", "impl", "expl0", "expl1", "t4", "t5", "t6", "t7"):
                        v = getattr(sy, f)
                        if v:
                            _inside(ext, v, v + 16 * L * nic, f"ic rhs {f}")
                for v in (a.rb, a.rj):
                    _inside(ext, v, v + 16 * L * N, "outer rhs")
                for v in (a.out_b, a.out_j):
                    _inside(ext, v, v + 16 * L * (N + nic), "coupled rhs")
                n_ic_calls += 1
            elif name == "mh_ic_impl":
                a = C.cast(args[0], C.POINTER(_lib.IcImplArgs)).contents
                for f in ("b", "db", "ddb", "aj", "dj", "ddj", "old_b", "old_j", "impl_b",
                          "impl_j"):
                    v = getattr(a, f)
                    if v:
                        _inside(ext, v, v + 16 * L * nic, f"ic impl {f}")
                _inside(ext, a.O_r_ic, a.O_r_ic + 8 * nic, "1/r inner")
                n_ic_calls += 1
    assert n_ic_calls == 2 * len(stages)


@pytest.mark.parametrize("scheme", ["CNAB2", "BPR353"])
def test_rotating_inner_core_launches_stay_inside_their_tensors(cpu_standin, scheme):
    """Rotating conducting inner core: the z(l=1, m=0) job and the z job without
    that row, the ICB synthesis / FFTs / analysis of the radial plan, and the
    arguments of mh_ic_torque, mh_ic_advect, mh_ic_omega_rhs and
    mh_ic_omega_update (rotation-state slots included) in every recorded plan
    lie inside the run's tensors."""
    cfg = dict(l_max=16, n_r_max=17, n_cheb_ic_max=7, sigma_ratio=1.0, kbotb=3, nRotIC=1,
               ek=1e-3, ra=1e5, pr=1.0, prmag=5.0, radratio=0.35, dtmax=1e-4, time_scheme=scheme)
    with _lib.record():
        run = dyn.DynamoRun(cfg, device="cpu")
        run.initialize_dt(cfg["dtmax"])
        run._rebuild_launches(0.5e-4)
    assert run.rot_ic
    ext = _extents(run)
    _check_row_jobs(run, ext)
    L, N, nic, T, P = run.L, run.N, run.p.n_r_ic_max, run.res.T, run.res.P
    # the z job skips (l=1, m=0), which has a job of its own
    jobs = np.frombuffer(run.jobs1[0].numpy().tobytes(), dtype=dyn._ROWJOB_DTYPE)
    z10 = jobs[jobs["rows"] == run.d_ic["lm10"].data_ptr()]
    assert len(z10) == 1 and z10["nrows"][0] == 1 and z10["MT"][0] == run.MT_z10.data_ptr()
    rows = run.res.d["l_rows"].numpy()
    zrows = [int(rows[(int(j["rows"]) - run.res.d["l_rows"].data_ptr()) // 4 + i])
             for j in jobs[jobs["X"] == run.rhs["z"].data_ptr()] if j["MT"] != z10["MT"][0]
             for i in range(int(j["nrows"]))]
    assert run.lm10 not in zrows and sorted(zrows) == sorted(set(range(L)) - {run.lm10})
    stages = (1, 2, 3, 4) if run.dirk else (0,)
    seen = set()
    for k in stages:
        run.stage = k
        kinds = ("radial", "lm") if k <= 3 else ("lm",)
        for kind in kinds:
            _, plan = run._plan(kind)
            for _, name, args in plan:
                if not name.startswith("mh_ic_") or name in ("mh_ic_rhs", "mh_ic_impl"):
                    continue
                a = C.cast(args[0], C.POINTER(_lib.IcRotArgs)).contents
                seen.add(name)
                _inside(ext, a.rot, a.rot + 8 * 10, "rotation state")
                assert 6 <= a.expl_slot <= 9 and 2 <= a.impl_slot <= 5
                assert all(-1 <= t <= 9 for t in a.term)
                assert (a.L, a.N, a.n_ic, a.T, a.P, a.lm10) == (L, N, nic, T, P, run.lm10)
                if name == "mh_ic_torque":
                    assert a.expl_slot == 6 + (max(0, k - 1) if run.dirk else run._par)
                    for v in (a.br, a.bp, a.vt, a.vp):
                        _inside(ext, v, v + 8 * T * P, "ICB grid")
                    for v in (a.gauss, a.sin2):
                        _inside(ext, v, v + 8 * T, "theta weights")
                elif name == "mh_ic_advect":
                    for v in (a.b_ic, a.aj_ic, a.e_b, a.e_j):
                        _inside(ext, v, v + 16 * L * nic, "inner field")
                    _inside(ext, a.m, a.m + 4 * L, "orders")
                    _inside(ext, a.dLh, a.dLh + 8 * L, "l(l+1)")
                elif name == "mh_ic_omega_rhs":
                    _inside(ext, a.rhs_z, a.rhs_z + 16 * L * N, "z rhs")
                    used = [t for t in a.term if t >= 0]
                    assert used[:2] == [1, 2]
                    if run.dirk:
                        assert len(used) == {1: 3, 2: 5, 3: 7, 4: 8}[k]
                else:
                    for v in (a.z, a.dz):
                        _inside(ext, v, v + 16 * L * N, "z")
                    assert a.store_old == (k == 4 if run.dirk else 1)
                    assert a.impl_slot == 2 + (k % 4 if run.dirk else 0)
            if kind == "radial":
                ana = [args for _, name, args in plan if name == "mh_leg_analysis"]
                a = C.cast(ana[-1][1], C.POINTER(_lib.AnaArgs)).contents
                assert (a.B, a.ldo, a.prec) == (1, N, 0)
                lo = a.outS[0]
                assert lo == run.nl["VxBt"].data_ptr() + 16 * (N - 1)
                _inside(ext, lo, lo + 16 * ((L - 1) * N + 1), "ICB VxBt")
    assert seen == {"mh_ic_torque", "mh_ic_advect", "mh_ic_omega_rhs", "mh_ic_omega_update"}


@pytest.mark.parametrize("cfg", [
    dict(l_max=12, n_r_max=17, mode=1, strat=3.0, polind=2.0, g2=1.0, ktopv=1, kbotv=1,
         l_correct_AMz=True, l_correct_AMe=True, dtmax=1e-4),
    dict(l_max=12, n_r_max=17, strat=1.0, polind=2.0, g1=1.0, dtmax=1e-4, minc=2),
    dict(l_max=12, n_r_max=17, strat=1.0, polind=2.0, g1=1.0, dtmax=1e-4,
         time_scheme="BPR353")])
def test_anelastic_launches_stay_inside_their_tensors(cpu_standin, cfg):
    """Anelastic step: every synthesis, Fourier derivative, FFT, grid-product,
    analysis, entropy-finish, angular-momentum, l = 0 pressure and implicit-term
    launch of the recorded radial and lm plans (each BPR353 stage) reads and
    writes inside the run's tensors."""
    with _lib.record():
        run = dyn.DynamoRun(cfg, device="cpu")
        run.initialize_dt(cfg["dtmax"])
        run._rebuild_launches(0.5e-4)
    assert run.anel and not run.fused_grid
    ext = _extents(run)
    _check_row_jobs(run, ext)
    r, N, L = run.res, run.N, run.L
    Nb, M, T, P = N - 2, r.M, r.T, r.P
    npt = T * P
    stages = (1, 2, 3, 4) if run.dirk else (0,)
    seen = set()
    for k in stages:
        run.stage = k
        for kind in (("radial", "lm") if k <= 3 else ("lm",)):
            _, plan = run._plan(kind)
            for _, name, args in plan:
                seen.add(name)
                if name == "mh_leg_synth":
                    a = C.cast(args[1], C.POINTER(_lib.SynArgs)).contents
                    nc = 3 if args[0] else 1
                    _inside(ext, a.F, a.F + 8 * (nc * a.c_stride), "synth F")
                    for g in range(a.ngroups):
                        for c in range(nc):
                            x = a.in_[g][c].x1
                            _inside(ext, x, x + 16 * L * a.ldx, "synth input")
                elif name == "mh_fft":
                    a = C.cast(args[1], C.POINTER(_lib.FftArgs)).contents
                    _inside(ext, a.F, a.F + 8 * a.ncomp * a.Fc_stride, "fft F")
                    _inside(ext, a.G, a.G + 8 * a.ncomp * a.Gc_stride, "fft G")
                elif name == "mh_fourier_dphi":
                    a = C.cast(args[0], C.POINTER(_lib.DphiArgs)).contents
                    _inside(ext, a.F, a.F + 8 * ((a.c0 + 2) * a.Fc_stride), "dphi F")
                    _inside(ext, a.out, a.out + 8 * 2 * a.Oc_stride, "dphi out")
                    assert a.B <= a.FB and a.B <= a.OB and a.n_m == M
                elif name == "mh_nl_anel":
                    a = C.cast(args[0], C.POINTER(_lib.NlAnelArgs)).contents
                    _inside(ext, a.Gs, a.Gs + 8 * 2 * N * npt, "nl s")
                    _inside(ext, a.Gv, a.Gv + 8 * 3 * a.Gv_c_stride, "nl u")
                    assert a.Gv_c_stride == 3 * N * npt and a.Gd_c_stride == N * npt
                    if a.l_mag:
                        _inside(ext, a.Gb, a.Gb + 8 * 3 * a.Gb_c_stride, "nl B")
                        assert a.Gb_c_stride == 2 * N * npt
                    _inside(ext, a.Gd, a.Gd + 8 * 2 * a.Gd_c_stride, "nl dphi")
                    _inside(ext, a.Go, a.Go + 8 * 10 * a.Go_c_stride, "nl out")
                    assert a.Go_c_stride == Nb * npt
                    _inside(ext, a.cfl, a.cfl + 8 * 2 * N, "cfl")
                    for v in (a.prof.beta, a.prof.orho1, a.prof.otemp1, a.prof.visc):
                        _inside(ext, v, v + 8 * N, "profile")
                    _inside(ext, a.cosn2, a.cosn2 + 8 * T, "cos/sin^2")
                elif name == "mh_leg_analysis":
                    a = C.cast(args[1], C.POINTER(_lib.AnaArgs)).contents
                    for i in range(a.ncomp):
                        lo = a.outS[i]
                        _inside(ext, lo, lo + 16 * ((L - 1) * a.ldo + a.B), "analysis out")
                elif name == "mh_entropy_anel":
                    a = C.cast(args[0], C.POINTER(_lib.EntropyAnelArgs)).contents
                    for v in (a.es, a.heat):
                        _inside(ext, v, v + 16 * L * N, "entropy")
                elif name == "mh_am_correct":
                    a = C.cast(args[0], C.POINTER(_lib.AmArgs)).contents
                    for v in (a.z, a.dz, a.d2z):
                        _inside(ext, v, v + 16 * L * N, "am z")
                    assert 0 <= a.l1m0 < L and -1 <= a.l1m1 < L
                elif name == "mh_p0_integral":
                    a = C.cast(args[0], C.POINTER(_lib.P0IntArgs)).contents
                    _inside(ext, a.rwp, a.rwp + 16 * L * 2 * N, "p0 rhs")
                    _inside(ext, a.u, a.u + 8 * N, "p0 weights")
                elif name == "mh_impl_anel":
                    a = C.cast(args[0], C.POINTER(_lib.ImplAnelArgs)).contents
                    for f in ("s", "z", "w", "dw", "ddw", "dddw", "p", "dp", "old_s", "impl_s",
                              "impl_w", "impl_p", "impl_z"):
                        v = getattr(a.base, f)
                        _inside(ext, v, v + 16 * L * N, f"impl {f}")
                    _inside(ext, a.dz, a.dz + 16 * L * N, "impl dz")
                    assert a.zero_edges == 0
    assert {"mh_nl_anel", "mh_fourier_dphi", "mh_entropy_anel", "mh_impl_anel"} <= seen
    if run.p.l_correct_AMz:
        assert "mh_am_correct" in seen
    if run.p0_integral:
        assert "mh_p0_integral" in seen


@pytest.mark.parametrize("cfg", [
    dict(l_max=12, n_r_max=33, radial_scheme="FD", fd_order=2, fd_order_bound=2, dtmax=1e-4),
    dict(l_max=12, n_r_max=40, radial_scheme="FD", fd_order=4, fd_order_bound=2, dtmax=1e-4,
         minc=2, time_scheme="BPR353"),
    dict(l_max=12, n_r_max=33, radial_scheme="FD", strat=0.5, polind=2.0, g2=1.0, mode=1,
         ktopv=1, kbotv=1, dtmax=1e-4),
    dict(l_max=12, n_r_max=33, radial_scheme="FD", sigma_ratio=1.0, kbotb=3, nRotIC=1,
         n_cheb_ic_max=15, dtmax=1e-4)])
def test_finite_difference_launches_stay_inside_their_tensors(cpu_standin, cfg):
    """Finite-difference (double-curl) step: the rebuild (affine assembly,
    inverses, folds, row fills), the row-operator jobs (including the second
    set-up pass over w''), and the double-curl explicit, buoyancy and
    implicit-term launches of every recorded plan read and write inside the
    run's tensors; the W row operator keeps its fixed l = 0 pressure block."""
    with _lib.record() as rec:
        run = dyn.DynamoRun(cfg, device="cpu")
        run.initialize_dt(cfg["dtmax"])
        run._rebuild_launches(0.5e-4)
    calls = rec.calls
    assert run.dc
    ext = _extents(run)
    _check_row_jobs(run, ext)
    dj, tile_lists = run.jobs_setup2
    assert len(np.frombuffer(dj.numpy().tobytes(), dtype=dyn._ROWJOB_DTYPE)) == 1
    N, L, lp1 = run.N, run.L, run.p.l_max + 1
    n = _check_rebuild_calls(calls, ext)
    aff = [a for _, name, a in calls if name == "mh_affine_f64"]
    assert n >= 6 and len(aff) == 2 + int(run.cond_ic) + int(run.rot_ic)
    for A0, A1, out, cnt in (a[:4] for a in aff):
        for v in (A0, A1, out):
            _inside(ext, v, v + 8 * cnt, "affine operand")
    nband = 0
    for _, name, a in calls:
        if name == "mh_band_factor_gated":
            abd, piv, inf, n, kl, ku, nb = a[:7]
            _inside(ext, abd, abd + 8 * nb * (2 * kl + ku + 1) * n, "band factor abd")
            _inside(ext, piv, piv + 4 * nb * n, "band pivots")
            _inside(ext, inf, inf + 4 * nb, "band info")
            nband += 1
        elif name == "mh_band_solve_gated":
            abd, piv, fr, fc, lm2l, rows, rhs, sol, nr, n, kl, ku = a[:12]
            _inside(ext, rhs, rhs + 16 * nr * n, "band rhs")
            _inside(ext, sol, sol + 16 * nr * n, "band solution")
            _inside(ext, lm2l, lm2l + 4 * nr, "band degrees")
            nl = int(np.frombuffer(C.string_at(lm2l, 4 * nr), np.int32).max()) + 1
            _inside(ext, abd, abd + 8 * nl * (2 * kl + ku + 1) * n, "band solve abd")
            _inside(ext, piv, piv + 4 * nl * n, "band solve pivots")
            nband += 1
    assert nband == 2 + run.nsys + 1
    mt0 = run.MT["wp"][0].cpu().numpy()
    assert np.abs(mt0[N:2 * N, 4 * N:5 * N]).max() > 0 and np.abs(mt0[:N]).max() == 0
    stages = (1, 2, 3, 4) if run.dirk else (0,)
    seen = set()
    for k in stages:
        run.stage = k
        for kind in (("radial", "lm") if k <= 3 else ("lm",)):
            _, plan = run._plan(kind)
            for _, name, args in plan:
                if name not in ("mh_td_dc", "mh_buoy_dc", "mh_impl_dc"):
                    continue
                seen.add(name)
                a = C.cast(args[0], C.POINTER(_lib.DcArgs)).contents
                assert (a.rc.N, a.lc.L) == (N, L)
                full = [a.w, a.dw, a.ddw, a.z, a.dz, a.s]
                if name == "mh_td_dc":
                    full += [a.Advr, a.dAdvt_dr, a.ew, a.ep]
                elif name == "mh_buoy_dc":
                    full += [a.ew0]
                    _inside(ext, a.rwp, a.rwp + 16 * L * 2 * N, "buoyancy rhs")
                else:
                    full += [a.dddw, a.ddddw, a.impl_w, a.old_w, a.old_p, a.impl_p]
                for v in full:
                    _inside(ext, v, v + 16 * L * N, name)
                _inside(ext, a.grav, a.grav + 8 * N, "gravity")
                if run.anel:
                    for v in (a.beta, a.ddbeta, a.orho1, a.rho0, a.visc, a.ddLvisc):
                        _inside(ext, v, v + 8 * N, "profile")
    assert seen == {"mh_td_dc", "mh_buoy_dc", "mh_impl_dc"}
