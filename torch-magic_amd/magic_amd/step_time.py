"""Step entry points with the reference's names (``step_time.py``).

``setup_initial_state``, ``initialize_dt``, ``one_step``, ``build_all_matrices``,
``radial_loop`` and ``lm_loop`` act on the default run of
:mod:`magic_amd.config` (reference ``step_time.py:398,1554,1574,1530,931,1478``).
"""

from __future__ import annotations

from . import config


def setup_initial_state():
    config.default_run().setup_initial_state()


def initialize_dt(dt: float):
    config.default_run().initialize_dt(dt)


def one_step(n_time_step: int, dt: float) -> float:
    return config.default_run().one_step(n_time_step, dt)


def build_all_matrices():
    run = config.default_run()
    run.build_all_matrices(run.p.alpha * run.dt_hist[0])


def radial_loop():
    return config.default_run().radial_loop()


def lm_loop():
    config.default_run().lm_loop()
