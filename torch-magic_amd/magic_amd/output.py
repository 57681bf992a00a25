"""Energy diagnostics and their log lines with the reference's names
(``output.py:76-80,111-193,304-320,486-501``)."""

from __future__ import annotations

from collections import namedtuple

from . import config

EKin = namedtuple("EKin", "e_p e_t e_p_as e_t_as e_p_es e_t_es e_p_eas e_t_eas")
EMagOC = namedtuple("EMagOC", "e_p e_t e_p_as e_t_as e_p_os e_p_as_os "
                              "e_p_es e_t_es e_p_eas e_t_eas e_p_e e_p_as_e")


def get_e_kin_full(w=None, dw=None, z=None):
    """The eight kinetic-energy components of the default run's current state
    (total, axisymmetric, equatorially symmetric and its m = 0 part, each
    poloidal and toroidal)."""
    return EKin(*config.default_run().energy_components()[0])


def get_e_mag_oc_full(b=None, db=None, aj=None):
    """The twelve outer-core magnetic-energy components of the default run
    (the external-field pair is zero, as in the reference's benchmarks)."""
    return EMagOC(*config.default_run().energy_components()[1])


def get_e_kin(w=None, dw=None, ddw=None, z=None, dz=None):
    """(e_kin_pol, e_kin_tor)."""
    ek = get_e_kin_full()
    return ek.e_p, ek.e_t


def get_e_mag(b=None, db=None, ddb=None, aj=None, dj=None):
    """(e_mag_pol, e_mag_tor)."""
    em = get_e_mag_oc_full()
    return em.e_p, em.e_t


def get_spectra():
    """Per-degree and per-order kinetic/magnetic energy spectra."""
    return config.default_run().spectra()


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
