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
