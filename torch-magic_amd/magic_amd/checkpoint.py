"""Restart files of a dynamo run in the reference's ``.pt`` layout.

Mirrors ``save_checkpoint`` / ``load_checkpoint`` of the reference
(``main.py:189-267``): one ``torch.save`` dict with ``step``, ``sim_time``,
``cfg``, ``tscheme_dt`` (dt of the last step, dt of the step before),
``fields.<name>`` (lm_max, n_r_max) complex arrays under the reference's field
names (``main.py:169-178``), ``fields.omega_ic`` / ``fields.omega_ma`` and, per
time array of ``dt_fields.py`` (``dsdt`` ... ``djdt``), ``dt.<name>.impl`` /
``.expl`` / ``.old`` of shape (lm_max, n_r_max, n) with the CNAB2 slot counts
nimp=1, nexp=2, nold=1 (``time_scheme.py:28-30``), lm in the same m-major order.

The file holds the state after a whole step, as the reference's main loop
writes it: ``expl[..., 1]`` is the explicit term of the last step and, since
the reference shifts its history at the end of a step (``rotate_imex``,
``time_scheme.py:130-143``), ``expl[..., 0]`` holds the same values.  Fields
this build does not evolve (composition ``xi``, the rotation rates of the
inner core and mantle) are written as zeros and must be zero on load.  Tensors
are stored on the CPU and read with ``torch.load(weights_only=True)``.
"""
from __future__ import annotations

import torch

from .params import PARAM_DEFAULTS

# reference field name -> key of DynamoRun.f (None: not evolved here, zero)
FIELD_NAMES = {
    "w_LMloc": "w", "dw_LMloc": "dw", "ddw_LMloc": "ddw",
    "z_LMloc": "z", "dz_LMloc": "dz",
    "p_LMloc": "p", "dp_LMloc": "dp",
    "s_LMloc": "s", "ds_LMloc": "ds",
    "xi_LMloc": None, "dxi_LMloc": None,
    "b_LMloc": "b", "db_LMloc": "db", "ddb_LMloc": "ddb",
    "aj_LMloc": "aj", "dj_LMloc": "dj", "ddj_LMloc": "ddj",
}
# reference time array -> system key of DynamoRun.old / impl / expl_*
DT_NAMES = {"dsdt": "s", "dxidt": None, "dwdt": "w", "dzdt": "z", "dpdt": "p",
            "dbdt": "b", "djdt": "j"}
SCALAR_DT_NAMES = ("domega_ic_dt", "domega_ma_dt")
NOLD, NEXP, NIMP = 1, 2, 1


def _cpu(t: torch.Tensor) -> torch.Tensor:
    return t.detach().to("cpu", copy=True)


def checkpoint_state(run, step: int | None = None, sim_time: float | None = None) -> dict:
    """The restart dict of `run` (after a whole step; not while device-controlled
    steps are queued)."""
    ctl = getattr(run, "_ctl", None)
    if ctl is not None and ctl.get("queued", run.n_step) > run.n_step:
        raise RuntimeError("checkpoint_state: call ctl_end() before writing a restart file")
    if isinstance(run.f["s"], torch.Tensor) and run.f["s"].is_cuda:
        torch.cuda.synchronize(run.f["s"].device)
    p = run.p
    cfg = {k: getattr(p, k) for k in PARAM_DEFAULTS}
    cfg.update(p.run)
    state = {
        "step": int(run.n_step if step is None else step),
        "sim_time": float(run.time if sim_time is None else sim_time),
        "cfg": cfg,
        "tscheme_dt": torch.tensor([float(run.dt_hist[0]), float(run.dt_hist[1])],
                                   dtype=torch.float64),
    }
    ref = run.f["s"]
    for name, key in FIELD_NAMES.items():
        state[f"fields.{name}"] = (_cpu(run.f[key]) if key is not None
                                   else torch.zeros(ref.shape, dtype=ref.dtype))
    state["fields.omega_ic"] = 0.0
    state["fields.omega_ma"] = 0.0
    L, N = ref.shape
    for name, key in DT_NAMES.items():
        if key is None or (key in ("b", "j") and not p.l_mag):
            z = lambda n: torch.zeros(L, N, n, dtype=ref.dtype)
            state[f"dt.{name}.impl"], state[f"dt.{name}.expl"], state[f"dt.{name}.old"] = \
                z(NIMP), z(NEXP), z(NOLD)
            continue
        last = _cpu(run.expl_prev[key])  # the step just taken (after the history swap)
        state[f"dt.{name}.impl"] = _cpu(run.impl[key]).unsqueeze(-1)
        state[f"dt.{name}.expl"] = torch.stack([last, last], dim=-1)
        state[f"dt.{name}.old"] = _cpu(run.old[key]).unsqueeze(-1)
    for name in SCALAR_DT_NAMES:
        state[f"dt.{name}.impl"] = torch.zeros(NIMP, dtype=torch.float64)
        state[f"dt.{name}.expl"] = torch.zeros(NEXP, dtype=torch.float64)
        state[f"dt.{name}.old"] = torch.zeros(NOLD, dtype=torch.float64)
    return state


def save_checkpoint(run, path, step: int | None = None, sim_time: float | None = None) -> None:
    """Write the restart file of `run` (reference ``main.save_checkpoint``)."""
    if getattr(run, "dirk", False):
        raise NotImplementedError("restart files hold CNAB2 time arrays; BPR353 runs are not "
                                  "written to them")
    if getattr(run, "cond_ic", False):
        raise NotImplementedError("restart files of runs with a conducting inner core are not "
                                  "written (the inner-core arrays are not stored)")
    torch.save(checkpoint_state(run, step, sim_time), path)


def _check_zero(state: dict, key: str) -> None:
    v = state.get(key)
    if v is None:
        return
    nz = bool(torch.count_nonzero(v)) if isinstance(v, torch.Tensor) else float(v) != 0.0
    if nz:
        raise ValueError(f"load_checkpoint: {key} is nonzero; this build does not evolve it")


def apply_checkpoint_state(run, state: dict) -> tuple[int, float]:
    """Restore `run` from a restart dict; returns (step, sim_time)."""
    if getattr(run, "dirk", False):
        raise NotImplementedError("restart files hold CNAB2 time arrays; BPR353 runs are not "
                                  "restarted from them")
    if getattr(run, "cond_ic", False):
        raise NotImplementedError("runs with a conducting inner core are not restarted from "
                                  "restart files")
    ref = run.f["s"]
    shape = tuple(ref.shape)
    for name, key in FIELD_NAMES.items():
        k = f"fields.{name}"
        if key is None:
            _check_zero(state, k)
            continue
        v = state[k]
        if tuple(v.shape) != shape:
            raise ValueError(f"load_checkpoint: {k} has shape {tuple(v.shape)}, run has {shape}")
        run.f[key].copy_(v.to(ref.device, ref.dtype))
    for k in ("fields.omega_ic", "fields.omega_ma"):
        _check_zero(state, k)
    for name in SCALAR_DT_NAMES:
        for part in ("impl", "expl", "old"):
            _check_zero(state, f"dt.{name}.{part}")
    for name, key in DT_NAMES.items():
        if key is None or (key in ("b", "j") and not run.p.l_mag):
            for part in ("impl", "expl", "old"):
                _check_zero(state, f"dt.{name}.{part}")
            continue
        impl, expl, old = (state[f"dt.{name}.{part}"] for part in ("impl", "expl", "old"))
        if (tuple(expl.shape) != shape + (NEXP,) or tuple(impl.shape) != shape + (NIMP,)
                or tuple(old.shape) != shape + (NOLD,)):
            raise ValueError(f"load_checkpoint: dt.{name} arrays do not fit a CNAB2 run of "
                             f"shape {shape}")
        run.impl[key].copy_(impl[..., 0].to(ref.device, ref.dtype))
        run.old[key].copy_(old[..., 0].to(ref.device, ref.dtype))
        # the next step overwrites the newest slot and reads slot 1 as its history
        run.expl_prev[key].copy_(expl[..., 1].to(ref.device, ref.dtype))
        run.expl_cur[key].copy_(expl[..., 0].to(ref.device, ref.dtype))
    dt = state["tscheme_dt"]
    run.dt_hist = [float(dt[0]), float(dt[1])]
    run.time = float(state["sim_time"])
    run.n_step = int(state["step"])
    # Courant cap: a run that never had initialize_dt (a fresh process) takes the
    # configured dtmax, clamped as initialize_dt clamps it.  The reference keeps
    # its cap at 0.0 there (step_time.py:1551), so its first decision after such
    # a restart would be dt = 0.
    if not getattr(run, "dtMax", None) and hasattr(run, "dt_cap"):
        run.dtMax = run.dt_cap(run.p.dtmax)
    # device step control restarts from the restored state: the next ctl_begin
    # pushes it (no queued or pulled steps of the state before the load remain)
    if getattr(run, "_ctl", None) is not None:
        run._ctl["pulled"] = run._ctl["queued"] = run.n_step
    if hasattr(run, "dt_log"):
        del run.dt_log[run.n_step:]
    # operators for the restored dt, as the reference rebuilds after a load
    run.build_all_matrices(run.p.alpha * run.dt_hist[0])
    return run.n_step, run.time


def load_checkpoint(run, path) -> tuple[int, float]:
    """Restore `run` from a restart file (reference ``main.load_checkpoint``);
    returns (step, sim_time).  Continue with ``run.one_step(step + 1, dt)``,
    dt = ``run.dt_hist[0]``."""
    return apply_checkpoint_state(run, torch.load(path, map_location="cpu", weights_only=True))
