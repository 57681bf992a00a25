"""Restart entry points under the reference's module name (``main.py:189-267``),
acting on the configured default run (``config.configure``)."""
from __future__ import annotations

from . import config
from .checkpoint import load_checkpoint as _load, save_checkpoint as _save


def save_checkpoint(path, step, sim_time, cfg=None):
    """Write the default run's restart file (reference ``main.save_checkpoint``;
    the file's ``cfg`` entry is the run's own configuration)."""
    _save(config.default_run(), path, step, sim_time)


def load_checkpoint(path):
    """Restore the default run from a restart file; returns (step, sim_time)
    (reference ``main.load_checkpoint``)."""
    return _load(config.default_run(), path)
