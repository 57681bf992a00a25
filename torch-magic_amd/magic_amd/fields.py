"""Spectral fields of the default run under the reference's names (``fields.py``).

``fields.s_LMloc`` etc. resolve to the default run's device tensors, (lm_max,
n_r_max) complex128 in m-major order; writes through them update the run.
"""

from __future__ import annotations

from . import config

_NAMES = {"s_LMloc": "s", "ds_LMloc": "ds", "z_LMloc": "z", "dz_LMloc": "dz", "w_LMloc": "w",
          "dw_LMloc": "dw", "ddw_LMloc": "ddw", "p_LMloc": "p", "dp_LMloc": "dp", "b_LMloc": "b",
          "db_LMloc": "db", "ddb_LMloc": "ddb", "aj_LMloc": "aj", "dj_LMloc": "dj",
          "ddj_LMloc": "ddj"}


def __getattr__(name):
    if name in _NAMES:
        return config.default_run().f[_NAMES[name]]
    raise AttributeError(name)
