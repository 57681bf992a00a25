"""Host logic of the restart files (no GPU): the reference's key set and
shapes (``main.py:169-223``, ``dt_fields.py``, CNAB2 slot counts of
``time_scheme.py:28-30``), a save/load round trip through torch.save with
weights_only loading, and refusal of state this build does not evolve."""

import types

import numpy as np
import pytest
import torch

from magic_amd import checkpoint as ck
from magic_amd.params import Params

REF_FIELDS = ["w_LMloc", "dw_LMloc", "ddw_LMloc", "z_LMloc", "dz_LMloc", "p_LMloc", "dp_LMloc",
              "s_LMloc", "ds_LMloc", "xi_LMloc", "dxi_LMloc", "b_LMloc", "db_LMloc",
              "ddb_LMloc", "aj_LMloc", "dj_LMloc", "ddj_LMloc"]
REF_DT = ["dsdt", "dxidt", "dwdt", "dzdt", "dpdt", "dbdt", "djdt"]


def _stub(L=10, N=7, seed=0):
    g = torch.Generator().manual_seed(seed)
    c = lambda: torch.complex(torch.randn(L, N, generator=g, dtype=torch.float64),
                              torch.randn(L, N, generator=g, dtype=torch.float64))
    sysk = ("s", "z", "w", "p", "b", "j")
    run = types.SimpleNamespace(
        p=Params({"l_max": 3, "n_r_max": N}, env=False),
        f={k: c() for k in ("s", "ds", "z", "dz", "w", "dw", "ddw", "p", "dp", "b", "db",
                            "ddb", "aj", "dj", "ddj")},
        old={k: c() for k in sysk}, impl={k: c() for k in sysk},
        expl_cur={k: c() for k in sysk}, expl_prev={k: c() for k in sysk},
        dt_hist=[2e-4, 3e-4], time=0.125, n_step=41, built=[])
    run.build_all_matrices = lambda w: run.built.append(w)
    return run


def test_restart_file_has_the_reference_layout(tmp_path):
    run = _stub()
    path = tmp_path / "r.pt"
    ck.save_checkpoint(run, path)
    st = torch.load(path, weights_only=True)
    expect = {"step", "sim_time", "cfg", "tscheme_dt", "fields.omega_ic", "fields.omega_ma"}
    expect |= {f"fields.{n}" for n in REF_FIELDS}
    expect |= {f"dt.{n}.{p}" for n in REF_DT + ["domega_ic_dt", "domega_ma_dt"]
               for p in ("impl", "expl", "old")}
    assert set(st) == expect
    assert st["step"] == 41 and st["sim_time"] == 0.125
    assert st["tscheme_dt"].tolist() == [2e-4, 3e-4]
    assert st["dt.dwdt.expl"].shape == (10, 7, 2) and st["dt.dwdt.old"].shape == (10, 7, 1)
    assert st["dt.domega_ic_dt.expl"].shape == (2,)
    # the step just taken sits in both explicit slots (history shifted at the step's end)
    assert torch.equal(st["dt.dwdt.expl"][..., 1], run.expl_prev["w"])
    assert torch.equal(st["dt.dwdt.expl"][..., 0], run.expl_prev["w"])
    assert torch.equal(st["fields.aj_LMloc"], run.f["aj"])
    assert st["cfg"]["l_max"] == 3


def test_restart_round_trip_restores_every_array(tmp_path):
    src, dst = _stub(seed=1), _stub(seed=2)
    path = tmp_path / "r.pt"
    ck.save_checkpoint(src, path)
    step, t = ck.load_checkpoint(dst, path)
    assert (step, t) == (41, 0.125) and dst.dt_hist == [2e-4, 3e-4]
    for k in src.f:
        assert torch.equal(dst.f[k], src.f[k]), k
    for d in ("old", "impl", "expl_prev"):
        for k in ("s", "z", "w", "p", "b", "j"):
            assert torch.equal(getattr(dst, d)[k], getattr(src, d)[k]), (d, k)
    # operators rebuilt for the restored dt (wimp_lin0 = alpha dt)
    assert dst.built == [src.p.alpha * 2e-4]


def test_restart_refuses_state_it_does_not_evolve(tmp_path):
    run = _stub()
    st = ck.checkpoint_state(run)
    st["fields.xi_LMloc"][0, 0] = 1.0
    with pytest.raises(ValueError, match="xi_LMloc"):
        ck.apply_checkpoint_state(_stub(), st)
    st = ck.checkpoint_state(run)
    st["fields.omega_ic"] = 0.5
    with pytest.raises(ValueError, match="omega_ic"):
        ck.apply_checkpoint_state(_stub(), st)
    st = ck.checkpoint_state(run)
    with pytest.raises(ValueError, match="shape"):
        ck.apply_checkpoint_state(_stub(L=12), st)


def _ic_stub(cfg):
    """A run-shaped object with the arrays of a conducting, rotating inner-core
    run of `cfg` (zeros)."""
    p = Params(cfg, env=False)
    L, N, nic = p.lm_max, p.n_r_max, p.n_r_ic_max
    dirk = p.time_scheme.upper() == "BPR353"
    z = lambda n=N: torch.zeros(L, n, dtype=torch.complex128)
    sysk = ("s", "z", "w", "p", "b", "j")
    d = lambda n=N: {k: z(n) for k in (sysk if n == N else ("b", "j"))}
    run = types.SimpleNamespace(
        p=p, dirk=dirk, cond_ic=True, rot_ic=True, _par=0,
        f={k: z() for k in ("s", "ds", "z", "dz", "w", "dw", "ddw", "p", "dp", "b", "db",
                            "ddb", "aj", "dj", "ddj")},
        fic={k: z(nic) for k in ck.IC_FIELD_NAMES},
        old=d(), impl=d(), expl_cur=d(), expl_prev=d(),
        old_ic=d(nic), impl_ic=d(nic), rot=torch.zeros(16, dtype=torch.float64),
        dt_hist=[0.0, 0.0], time=0.0, n_step=0, built=[])
    run.implS = [run.impl, d(), d(), d()] if dirk else [run.impl]
    run.explS = [run.expl_cur, run.expl_prev, d(), d()] if dirk else None
    run.implS_ic = [run.impl_ic, d(nic), d(nic), d(nic)] if dirk else [run.impl_ic]
    run.expl_ic = [d(nic) for _ in range(4 if dirk else 2)]
    run.build_all_matrices = lambda w: run.built.append(w)
    return run


@pytest.mark.parametrize("name", ["filesrot8", "filesbprrot8"])
def test_inner_core_restart_file_round_trips_the_reference_file(name):
    """The reference's .pt restart file of a conducting, rotating inner core
    (CNAB2 and BPR353, written after 10 steps): loaded into a run and written
    again it gives the same keys, shapes and values (CNAB2: the reference keeps
    the step's explicit term in both slots; BPR353: all four slots)."""
    import json
    import os
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    st = torch.load(os.path.join(gold, f"{name}_restart.pt"), weights_only=True)
    cfg = json.load(open(os.path.join(gold, f"{name}.json")))["cfg"]
    run = _ic_stub(cfg)
    step, t = ck.apply_checkpoint_state(run, st)
    assert (step, t) == (st["step"], st["sim_time"])
    assert run.rot[0] == st["fields.omega_ic"] != 0.0
    dirk = run.dirk
    assert run.built == [(0.5 if dirk else run.p.alpha) * float(st["tscheme_dt"][0])]
    out = ck.checkpoint_state(run)
    assert set(out) == set(st)
    for k, v in st.items():
        if k == "cfg":
            continue
        if isinstance(v, torch.Tensor):
            assert out[k].shape == v.shape and torch.equal(out[k], v), k
        else:
            assert out[k] == v, k
