"""MagIC Fortran stream checkpoints (reference ``checkpoint_io.py``): the
reader against a file the reference's own writer produced, and the writer
against that file byte for byte (tests/golden/files16_checkpoint.bin, written
after 10 steps of an l_max=16, n_r_max=17 run, with the state it was written
from in files16.npz)."""

import os
import types

import numpy as np
import pytest
import torch

from magic_amd import checkpoint_io as cio
from magic_amd import grid
from magic_amd.params import Params

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CFG = dict(l_max=16, n_r_max=17, ek=1e-3, ra=1e5, pr=1.0, prmag=5.0, radratio=0.35,
           dtmax=1e-4)


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "files16.npz"))


def test_reader_reads_the_reference_file(gold):
    ck = cio.read_checkpoint(os.path.join(GOLD, "files16_checkpoint.bin"))
    assert ck.version == 5 and ck.endian == ">" and ck.family == "MULTISTEP"
    assert (ck.nexp, ck.nimp, ck.nold) == (2, 1, 1) and ck.n_time_step == 11
    assert ck.time == float(gold["time"]) and np.array_equal(ck.dt, gold["tscheme_dt"])
    assert (ck.l_max, ck.m_max, ck.minc, ck.n_r_max, ck.n_theta_max, ck.n_phi_tot,
            ck.n_r_ic_max) == (16, 16, 1, 17, 24, 48, 9)
    assert ck.l_heat and ck.l_mag and ck.l_press_store and not ck.l_cond_ic
    assert (ck.ra, ck.ek, ck.prmag, ck.radratio) == (1e5, 1e-3, 5.0, 0.35)
    for k in ("w", "z", "p", "s", "b", "aj"):
        a = getattr(ck, k)
        assert a.dtype == np.complex128 and a.dtype.isnative
        assert np.array_equal(a, gold[k]), k
        assert len(ck.derivatives[k]["expl"]) == 1 and not ck.derivatives[k]["impl"]
        assert np.array_equal(ck.derivatives[k]["expl"][0], gold["expl1_" + k]), k
    assert np.abs(ck.r - grid.radial(Params(CFG))["r"]).max() <= 1e-15 * ck.r[0]


def test_reader_refuses_truncated_and_padded_files(tmp_path):
    raw = open(os.path.join(GOLD, "files16_checkpoint.bin"), "rb").read()
    (tmp_path / "short.bin").write_bytes(raw[:-8])
    (tmp_path / "long.bin").write_bytes(raw + b"\0" * 8)
    (tmp_path / "bad.bin").write_bytes(b"\x07\x07\x07\x07" + raw[4:])
    for name in ("short.bin", "long.bin", "bad.bin"):
        with pytest.raises(ValueError):
            cio.read_checkpoint(tmp_path / name)


def _state_from_gold(gold, r):
    p = Params(CFG)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    sysk = dict(w="w", z="z", p="p", s="s", b="b", aj="j")
    return types.SimpleNamespace(
        p=p, res=types.SimpleNamespace(rad={"r": r}),
        f={k: t(gold[k]) for k in ("w", "z", "p", "s", "b", "aj")},
        expl_prev={sysk[k]: t(gold["expl1_" + k]) for k in sysk},
        expl_cur={sysk[k]: t(np.zeros_like(gold[k])) for k in sysk},
        expl_pending=False, dt_hist=[float(x) for x in gold["tscheme_dt"]])


@pytest.mark.parametrize("endian", [">", "<"])
def test_writer_reproduces_the_reference_file(gold, tmp_path, endian):
    """Same state -> same bytes (big-endian, as the reference writes); the
    little-endian variant reads back to the same values."""
    ref = open(os.path.join(GOLD, "files16_checkpoint.bin"), "rb").read()
    r_ref = cio.read_checkpoint(os.path.join(GOLD, "files16_checkpoint.bin")).r
    run = _state_from_gold(gold, r_ref)
    out = tmp_path / "ck.bin"
    cio.write_checkpoint_fortran(out, run, float(gold["time"]), 11, endian=endian)
    if endian == ">":
        assert out.read_bytes() == ref
    ck = cio.read_checkpoint(out)
    assert ck.endian == endian
    for k in ("w", "s", "aj"):
        assert np.array_equal(getattr(ck, k), gold[k])
    # with this build's own radial grid (equal to the reference's within an ulp)
    run2 = _state_from_gold(gold, grid.radial(Params(CFG))["r"])
    cio.write_checkpoint_fortran(tmp_path / "ck2.bin", run2, float(gold["time"]), 11)
    ck2 = cio.read_checkpoint(tmp_path / "ck2.bin")
    assert np.abs(ck2.r - r_ref).max() <= 1e-15 * r_ref[0]
    assert np.array_equal(ck2.w, gold["w"])


IC_CASES = {"filesrot8": "MULTISTEP", "filesbprrot8": "DIRK"}


def _ic_state_from_file(name, ck):
    """A run-shaped state holding what the reference's file was written from:
    fields and inner fields from the file, explicit histories (MULTISTEP) from
    the file, the omega_ic history in the rotation buffer."""
    import json
    meta = json.load(open(os.path.join(GOLD, f"{name}.json")))
    p = Params(meta["cfg"])
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    dirk = ck.family == "DIRK"
    sysk = dict(w="w", z="z", p="p", s="s", b="b", aj="j")
    zero = lambda a: torch.zeros(a.shape, dtype=torch.complex128)
    hist = lambda k: t(ck.derivatives[k]["expl"][0]) if not dirk else zero(getattr(ck, k))
    rot = torch.zeros(16, dtype=torch.float64)
    if not dirk:
        rot[7] = ck.domega_ic_dt["expl"][0]   # parity 0: slot 7 holds the history
    return types.SimpleNamespace(
        p=p, res=types.SimpleNamespace(rad={"r": ck.r}), dirk=dirk, cond_ic=True, _par=0,
        f={k: t(getattr(ck, k)) for k in sysk},
        fic={k: t(getattr(ck, k)) for k in ("b_ic", "aj_ic")},
        expl_prev={sysk[k]: hist(k) for k in sysk},
        expl_cur={sysk[k]: zero(getattr(ck, k)) for k in sysk},
        expl_ic=[{"b": zero(ck.b_ic), "j": zero(ck.aj_ic)},
                 {"b": hist("b_ic"), "j": hist("aj_ic")}],
        rot=rot, expl_pending=False, dt_hist=[float(ck.dt[0]), float(ck.dt[-1])])


@pytest.mark.parametrize("name", sorted(IC_CASES))
def test_inner_core_checkpoint_reads_and_writes_like_the_reference(name, tmp_path):
    """Conducting, rotating inner core (reference files after 10 steps of CNAB2
    and of BPR353): the reader finds the family, slot counts, the inner fields
    and (MULTISTEP) the inner-core and omega_ic histories; the writer, given the
    same state, reproduces the file byte for byte."""
    g = np.load(os.path.join(GOLD, f"{name}.npz"))
    path = os.path.join(GOLD, f"{name}_checkpoint.bin")
    ck = cio.read_checkpoint(path)
    assert ck.family == IC_CASES[name] and ck.l_cond_ic and ck.l_mag
    assert (ck.nexp, ck.nimp, ck.nold) == ((4, 4, 1) if ck.family == "DIRK" else (2, 1, 1))
    assert ck.n_r_ic_max == 9 and ck.sigma_ratio == 1.0 and ck.time == float(g["time"])
    for k in ("w", "z", "s", "b", "aj", "b_ic", "aj_ic"):
        assert np.array_equal(getattr(ck, k), g[k]), k
    if ck.family == "MULTISTEP":
        assert set(ck.derivatives) == {"w", "z", "p", "s", "b", "aj", "b_ic", "aj_ic"}
        assert len(ck.domega_ic_dt["expl"]) == 1 and ck.domega_ic_dt["expl"][0] != 0.0
        assert np.any(ck.derivatives["b_ic"]["expl"][0])
    else:
        assert not ck.derivatives and ck.domega_ic_dt is None
    run = _ic_state_from_file(name, ck)
    out = tmp_path / "ck.bin"
    cio.write_checkpoint_fortran(out, run, ck.time, ck.n_time_step)
    assert out.read_bytes() == open(path, "rb").read()
