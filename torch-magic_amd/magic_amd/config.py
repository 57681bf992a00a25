"""Configuration surface of the reference: ``configure(dict)`` then use the modules.

Mirrors ``config.configure`` (reference ``config.py:19-33``): the dict (YAML
keys, snake-case Fortran namelist names) selects the run; environment
variables ``MAGIC_<KEY>`` are the fallback.  Unlike the reference, state is
not frozen at import time: the module-level facades (``step_time``,
``fields``, ``dt_fields``, ``sht``, ``output``) bind to a default
:class:`~magic_amd.dynamo.DynamoRun` created on first use from the current
configuration, and ``configure`` may be called again to start a new one.
"""

from __future__ import annotations

_config: dict = {}
_run = None
_res = None


def configure(overrides: dict | None = None):
    """Set the configuration of the default run (drops any existing default run)."""
    global _config, _run, _res
    _config = dict(overrides or {})
    _run = None
    _res = None
    from . import sht
    sht._default = None


def params():
    from .params import Params
    return Params(_config)


def default_run():
    """The default :class:`DynamoRun` of the configured run (created lazily)."""
    global _run, _res
    if _run is None:
        from .dynamo import DynamoRun
        p = params()
        _run = DynamoRun(p, device=p.device)
        _res = _run.res
    return _run


def default_resolution():
    global _res
    if _res is None:
        if _run is not None:
            _res = _run.res
        else:
            from .device import Resolution
            p = params()
            _res = Resolution(p, p.device)
    return _res
