"""Energy diagnostics with the reference's names (``output.py:111-193``)."""

from __future__ import annotations

from collections import namedtuple

from . import config

EKin = namedtuple("EKin", "e_p e_t")
EMagOC = namedtuple("EMagOC", "e_p e_t")


def get_e_kin_full(w=None, dw=None, z=None):
    """Kinetic poloidal/toroidal energies of the default run's current state."""
    e = config.default_run().energies()
    return EKin(e[0], e[1])


def get_e_mag_oc_full(b=None, db=None, aj=None):
    """Outer-core magnetic poloidal/toroidal energies of the default run."""
    e = config.default_run().energies()
    return EMagOC(e[2], e[3])


def get_spectra():
    """Per-degree and per-order kinetic/magnetic energy spectra."""
    return config.default_run().spectra()
