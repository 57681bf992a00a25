"""IMEX history of the default run under the reference's names (``dt_fields.py``).

``dt_fields.dsdt.old`` / ``.impl`` / ``.expl`` expose (lm_max, n_r_max, nslot)
views shaped like the reference's ``TimeArray`` (CNAB2: nold=1, nimp=1,
nexp=2); ``expl[..., 0]`` is the most recent explicit term.
"""

from __future__ import annotations

import torch

from . import config

_KEYS = {"dsdt": "s", "dzdt": "z", "dwdt": "w", "dpdt": "p", "dbdt": "b", "djdt": "j"}


class TimeArrayView:
    def __init__(self, run, key):
        self._run, self._k = run, key

    @property
    def old(self):
        return self._run.old[self._k].unsqueeze(-1)

    @property
    def impl(self):
        return self._run.impl[self._k].unsqueeze(-1)

    @property
    def expl(self):
        r = self._run
        # after a step the most recent explicit term is in the "previous" slot
        return torch.stack([r.expl_prev[self._k], r.expl_cur[self._k]], dim=-1)


def __getattr__(name):
    if name in _KEYS:
        return TimeArrayView(config.default_run(), _KEYS[name])
    raise AttributeError(name)
