"""Energy diagnostics and their log lines with the reference's names
(``output.py:76-80,101-193,304-320,486-501,750-861``).

Every function computes on the tensors it is given, as the reference's do:
(lm_max, n_r_max) complex arrays of the configured resolution, on any device
(moved to the GPU for the reduction), reduced by the device kernel
``mh_energy``.  The default run supplies the resolution (radial grid,
quadrature weights, lm maps)."""

from __future__ import annotations

from collections import namedtuple

from . import config

EKin = namedtuple("EKin", "e_p e_t e_p_as e_t_as e_p_es e_t_es e_p_eas e_t_eas")
EMagOC = namedtuple("EMagOC", "e_p e_t e_p_as e_t_as e_p_os e_p_as_os "
                              "e_p_es e_t_es e_p_eas e_t_eas e_p_e e_p_as_e")


def get_e_kin_full(w, dw, z):
    """The eight kinetic-energy components of (w, dw, z) (reference
    ``output.py:111-139``): total, axisymmetric, equatorially symmetric and its
    m = 0 part, each poloidal and toroidal."""
    run = config.default_run()
    return EKin(*run.energy_components(run.field_set(w=w, dw=dw, z=z))[0])


def get_e_mag_oc_full(b, db, aj):
    """The twelve outer-core magnetic-energy components of (b, db, aj)
    (reference ``output.py:152-193``; the external-field pair is zero there)."""
    run = config.default_run()
    if not run.p.l_mag:
        # the reference's formulas need no flag: run them on the field given
        raise ValueError("get_e_mag_oc_full: the configured run has no magnetic field "
                         "(mode 1 or 7)")
    return EMagOC(*run.energy_components(run.field_set(b=b, db=db, aj=aj))[1])


def get_e_kin(w, dw, ddw, z, dz):
    """(e_kin_pol, e_kin_tor) of the given fields (reference ``output.py:101-108``)."""
    ek = get_e_kin_full(w, dw, z)
    return ek.e_p, ek.e_t


def get_e_mag(b, db, ddb, aj, dj):
    """(e_mag_pol, e_mag_tor) of the given fields (reference ``output.py:142-149``)."""
    em = get_e_mag_oc_full(b, db, aj)
    return em.e_p, em.e_t


def get_dlm(w, dw, z, switch="V"):
    """Integral length scales (dl, dlR, dm, dlc, dlPolPeak, dlRc, dlPolPeakR)
    of the flow (switch 'V', arguments w, dw, z) or of the field ('B', arguments
    b, db, aj) (reference ``output.py:750-861``); dlR, dlRc and dlPolPeakR are
    (n_r_max,) float64 tensors, the rest floats."""
    import torch
    run = config.default_run()
    fs = run.field_set(w=w, dw=dw, z=z) if switch == "V" else run.field_set(b=w, db=dw, aj=z)
    dl, dlR, dm, dlc, pk, dlRc, pkR = run.get_dlm(switch, fs)
    t = lambda a: torch.as_tensor(a, dtype=torch.float64)
    return dl, t(dlR), dm, dlc, pk, t(dlRc), t(pkR)


def get_spectra(w=None, dw=None, z=None, b=None, db=None, aj=None):
    """Per-degree and per-order kinetic/magnetic energy spectra of the given
    fields (the default run's own where an argument is None)."""
    run = config.default_run()
    return run.spectra(run.field_set(w=w, dw=dw, z=z, b=b, db=db, aj=aj))


# --- log lines in the reference's Fortran formats (a file format: kept alike) ---

def write_e_kin_line(f, time, ek: EKin):
    """One e_kin.TAG line: ES20.12, 8ES16.8."""
    f.write(f"{time:20.12E}" + "".join(f"{v:16.8E}" for v in ek) + "\n")


def write_e_mag_oc_line(f, time, em: EMagOC):
    """One e_mag_oc.TAG line: ES20.12, 12ES16.8."""
    f.write(f"{time:20.12E}" + "".join(f"{v:16.8E}" for v in em) + "\n")


def write_timestep_line(f, time, dt):
    """One timestep.TAG line: ES20.12, ES16.8."""
    f.write(f"{time:20.12E}{dt:16.8E}\n")


def write_radius_file(path, r):
    """radius.TAG: (I4, ES16.8) per radial level."""
    with open(path, "w") as f:
        for i, v in enumerate(r):
            f.write(f"{i + 1:4d}{float(v):16.8E}\n")


def write_signal_file(path):
    """signal.TAG: the word NOT (no signal pending)."""
    with open(path, "w") as f:
        f.write("NOT\n")


class LogFiles:
    """The e_kin / e_mag_oc / timestep log files of one run, ``<name>.<tag>`` in
    `directory`, with radius.TAG and signal.TAG written at creation (the files
    the reference's main loop writes every log step).  ``log(time, dt)`` appends
    one line to each from the run's current state."""

    def __init__(self, run, directory, tag):
        import os
        self.run = run
        self.paths = {n: os.path.join(directory, f"{n}.{tag}")
                      for n in ("e_kin", "e_mag_oc", "timestep")}
        write_radius_file(os.path.join(directory, f"radius.{tag}"), run.res.rad["r"])
        write_signal_file(os.path.join(directory, f"signal.{tag}"))

    def log(self, time, dt):
        ek, em = self.run.energy_components()
        with open(self.paths["e_kin"], "a") as f:
            write_e_kin_line(f, time, EKin(*ek))
        if self.run.p.l_mag:
            with open(self.paths["e_mag_oc"], "a") as f:
                write_e_mag_oc_line(f, time, EMagOC(*em))
        with open(self.paths["timestep"], "a") as f:
            write_timestep_line(f, time, dt)
