"""IMEX history of the default run under the reference's names (``dt_fields.py``).

``dt_fields.dsdt.old`` / ``.impl`` / ``.expl`` stand for the reference's
``TimeArray`` arrays of shape (lm_max, n_r_max, nslot) (CNAB2: nold=1, nimp=1,
nexp=2; ``dt_fields.py:55-123``).  ``old`` and ``impl`` are views of the run's
device tensors.  The two explicit slots live in two buffers that swap roles
every step (the history shift of ``rotate_imex``, ``time_scheme.py:130-143``,
is a pointer swap here), so ``expl`` is a slot view: ``expl[..., k]`` (or
``expl[:, :, k]``) is the run's buffer of slot k itself, and assignments
through it (``expl[:, :, 1] = x``, ``expl[..., 1].copy_(y)``) update the run.

Slots as in the reference: between ``radial_loop`` and ``lm_loop`` slot 0
holds the new explicit term and slot 1 the previous one; after a whole step
(``rotate_imex`` copies slot 0 into slot 1) both read the newest term.  Here
slot 1 is always the history buffer the next step reads, and slot 0 is the
buffer of the new term while it awaits the lm loop and the history buffer
otherwise (the reference's slot 0 then holds a copy of it that the next
radial loop overwrites).
"""

from __future__ import annotations

import torch

from . import config

_KEYS = {"dsdt": "s", "dzdt": "z", "dwdt": "w", "dpdt": "p", "dbdt": "b", "djdt": "j"}


class ExplicitSlots:
    """(lm_max, n_r_max, 2) explicit-term history of one field: indexing whose
    last subscript is a slot number returns (a view of) that slot's device
    buffer; any other read returns a stacked copy."""

    def __init__(self, run, key):
        self._run, self._k = run, key

    def _slots(self):
        r = self._run
        hist = r.expl_prev[self._k]
        new = r.expl_cur[self._k] if r.expl_pending else hist
        return (new, hist)

    @property
    def shape(self):
        a = self._slots()[0]
        return torch.Size(tuple(a.shape) + (2,))

    @property
    def dtype(self):
        return self._slots()[0].dtype

    @property
    def device(self):
        return self._slots()[0].device

    def _split(self, idx):
        if not isinstance(idx, tuple):
            idx = (idx,)
        if Ellipsis in idx:
            i = idx.index(Ellipsis)
            fill = 3 - (len(idx) - 1)
            idx = idx[:i] + (slice(None),) * fill + idx[i + 1:]
        if len(idx) == 3 and isinstance(idx[2], int):
            k = idx[2]
            if not -2 <= k < 2:
                raise IndexError(f"explicit slot {k} out of range for 2 slots")
            return idx[:2], k % 2
        return None, idx

    def __getitem__(self, idx):
        head, k = self._split(idx)
        if head is not None:
            return self._slots()[k][head]
        return torch.stack(self._slots(), dim=-1)[k]

    def __setitem__(self, idx, value):
        head, k = self._split(idx)
        if head is None:
            raise IndexError("assign to one explicit slot at a time: expl[..., k] = value")
        self._slots()[k][head] = value

    def __array__(self, dtype=None):
        a = torch.stack(self._slots(), dim=-1).cpu().numpy()
        return a if dtype is None else a.astype(dtype)


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
        return ExplicitSlots(self._run, self._k)


def __getattr__(name):
    if name in _KEYS:
        return TimeArrayView(config.default_run(), _KEYS[name])
    raise AttributeError(name)
