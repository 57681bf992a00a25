"""Spectral fields of the default run under the reference's names (``fields.py``).

``fields.s_LMloc`` etc. resolve to the default run's device tensors, (lm_max,
n_r_max) complex128 in m-major order; writes through them update the run.  An
assignment to a name (``fields.s_LMloc = x``, or ``fields.s_LMloc += x``,
which the reference's callers use to add seeded noise) copies the value into
the run's tensor, so the name keeps resolving to the run's state.
"""

from __future__ import annotations

import sys
import types

from . import config

_NAMES = {"s_LMloc": "s", "ds_LMloc": "ds", "z_LMloc": "z", "dz_LMloc": "dz", "w_LMloc": "w",
          "dw_LMloc": "dw", "ddw_LMloc": "ddw", "p_LMloc": "p", "dp_LMloc": "dp", "b_LMloc": "b",
          "db_LMloc": "db", "ddb_LMloc": "ddb", "aj_LMloc": "aj", "dj_LMloc": "dj",
          "ddj_LMloc": "ddj"}


class _FieldsModule(types.ModuleType):
    def __getattr__(self, name):
        if name in _NAMES:
            return config.default_run().f[_NAMES[name]]
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name in _NAMES:
            t = config.default_run().f[_NAMES[name]]
            if value is not t:
                import torch
                t.copy_(torch.as_tensor(value).to(t.device, t.dtype))
            return
        super().__setattr__(name, value)


sys.modules[__name__].__class__ = _FieldsModule
