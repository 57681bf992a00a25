"""Restart files of a dynamo run in the reference's ``.pt`` layout.

Mirrors ``save_checkpoint`` / ``load_checkpoint`` of the reference
(``main.py:189-267``): one ``torch.save`` dict with ``step``, ``sim_time``,
``cfg``, ``tscheme_dt`` (CNAB2: dt of the last step and of the step before;
BPR353: dt), ``fields.<name>`` (lm_max, n_r_max) complex arrays under the
reference's field names (``main.py:169-178``), with a conducting inner core
also ``fields.b_ic`` ... ``fields.ddj_ic`` (lm_max, n_r_ic_max)
(``main.py:180-183``), ``fields.omega_ic`` / ``fields.omega_ma`` and, per time
array of ``dt_fields.py`` (``dsdt`` ... ``djdt``, ``dbdt_ic`` / ``djdt_ic``),
``dt.<name>.impl`` / ``.expl`` / ``.old`` of shape (lm_max, n_r, n) with the
scheme's slot counts (CNAB2 nimp=1, nexp=2, nold=1, ``time_scheme.py:28-30``;
BPR353 nimp=4, nexp=4, nold=1, ``time_scheme.py:176-181``), and per time scalar
(``domega_ic_dt``, ``domega_ma_dt``) the same three vectors; lm in the same
m-major order.

The file holds the state after a whole step, as the reference's main loop
writes it.  CNAB2: ``expl[..., 1]`` is the explicit term of the last step and,
since the reference shifts its history at the end of a step (``rotate_imex``,
``time_scheme.py:130-143``), ``expl[..., 0]`` holds the same values.  BPR353:
every slot as the last step left it (the next step reads only ``old`` and
``impl[..., 0]``).  Fields this build does not evolve (composition ``xi``, the
mantle's rotation, the inner core's rotation unless ``nRotIC = 1``) are written
as zeros and must be zero on load.  Tensors are stored on the CPU and read with
``torch.load(weights_only=True)``.
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
# reference inner-core time array -> system key of DynamoRun.old_ic / impl_ic / expl_ic
IC_DT_NAMES = {"dbdt_ic": "b", "djdt_ic": "j"}
# reference inner-core field names (keys of DynamoRun.fic)
IC_FIELD_NAMES = ("b_ic", "db_ic", "ddb_ic", "aj_ic", "dj_ic", "ddj_ic")
SCALAR_DT_NAMES = ("domega_ic_dt", "domega_ma_dt")
NOLD, NEXP, NIMP = 1, 2, 1              # CNAB2
DIRK_NOLD, DIRK_NEXP, DIRK_NIMP = 1, 4, 4  # BPR353


def slot_counts(run) -> tuple[int, int, int]:
    """(nold, nexp, nimp) of the run's time scheme."""
    return (DIRK_NOLD, DIRK_NEXP, DIRK_NIMP) if getattr(run, "dirk", False) else (NOLD, NEXP, NIMP)


def _cpu(t: torch.Tensor) -> torch.Tensor:
    return t.detach().to("cpu", copy=True)


def _slots(run, key, inner=False):
    """(old, impl slots, expl slots) of system `key` (outer arrays, or the inner
    core's) in the reference's slot order; CNAB2 lists the newest explicit term
    first, then the history the next step reads."""
    if inner:
        old = run.old_ic[key]
        impl = [d[key] for d in run.implS_ic] if run.dirk else [run.impl_ic[key]]
        E = run.expl_ic
        expl = ([d[key] for d in E] if run.dirk
                else [E[run._par][key], E[1 - run._par][key]])
        return old, impl, expl
    if getattr(run, "dirk", False):
        return run.old[key], [d[key] for d in run.implS], [d[key] for d in run.explS]
    return run.old[key], [run.impl[key]], [run.expl_cur[key], run.expl_prev[key]]


def _rot_slots(run):
    """(old, impl, expl) index lists of the rotation-state buffer, as _slots."""
    if run.dirk:
        return [1], [2, 3, 4, 5], [6, 7, 8, 9]
    return [1], [2], [6 + run._par, 7 - run._par]


def checkpoint_state(run, step: int | None = None, sim_time: float | None = None) -> dict:
    """The restart dict of `run` (after a whole step; not while device-controlled
    steps are queued)."""
    ctl = getattr(run, "_ctl", None)
    if ctl is not None and ctl.get("queued", run.n_step) > run.n_step:
        raise RuntimeError("checkpoint_state: call ctl_end() before writing a restart file")
    if isinstance(run.f["s"], torch.Tensor) and run.f["s"].is_cuda:
        torch.cuda.synchronize(run.f["s"].device)
    p = run.p
    dirk = getattr(run, "dirk", False)
    cond_ic = getattr(run, "cond_ic", False)
    nold, nexp, nimp = slot_counts(run)
    cfg = {k: getattr(p, k) for k in PARAM_DEFAULTS}
    cfg.update(p.run)
    dts = [float(run.dt_hist[0])] if dirk else [float(run.dt_hist[0]), float(run.dt_hist[1])]
    state = {
        "step": int(run.n_step if step is None else step),
        "sim_time": float(run.time if sim_time is None else sim_time),
        "cfg": cfg,
        "tscheme_dt": torch.tensor(dts, dtype=torch.float64),
    }
    ref = run.f["s"]
    for name, key in FIELD_NAMES.items():
        state[f"fields.{name}"] = (_cpu(run.f[key]) if key is not None
                                   else torch.zeros(ref.shape, dtype=ref.dtype))
    if cond_ic:
        for name in IC_FIELD_NAMES:
            state[f"fields.{name}"] = _cpu(run.fic[name])
    rot = _cpu(run.rot) if cond_ic else None
    state["fields.omega_ic"] = float(rot[0]) if cond_ic else 0.0
    state["fields.omega_ma"] = 0.0
    L, N = ref.shape

    def arrays(old, impl, expl):
        if not dirk:   # the step just taken in both explicit slots (shifted history)
            expl = [expl[1], expl[1]]
        return (torch.stack([_cpu(t) for t in impl], dim=-1),
                torch.stack([_cpu(t) for t in expl], dim=-1), _cpu(old).unsqueeze(-1))

    for name, key in DT_NAMES.items():
        if key is None or (key in ("b", "j") and not p.l_mag):
            z = lambda n: torch.zeros(L, N, n, dtype=ref.dtype)
            state[f"dt.{name}.impl"], state[f"dt.{name}.expl"], state[f"dt.{name}.old"] = \
                z(nimp), z(nexp), z(nold)
            continue
        state[f"dt.{name}.impl"], state[f"dt.{name}.expl"], state[f"dt.{name}.old"] = \
            arrays(*_slots(run, key))
    if cond_ic:
        for name, key in IC_DT_NAMES.items():
            state[f"dt.{name}.impl"], state[f"dt.{name}.expl"], state[f"dt.{name}.old"] = \
                arrays(*_slots(run, key, inner=True))
    for name in SCALAR_DT_NAMES:
        for part, n in (("impl", nimp), ("expl", nexp), ("old", nold)):
            state[f"dt.{name}.{part}"] = torch.zeros(n, dtype=torch.float64)
    if cond_ic:
        o, i, e = _rot_slots(run)
        if not dirk:
            e = [e[1], e[1]]
        for part, idx in (("impl", i), ("expl", e), ("old", o)):
            state[f"dt.domega_ic_dt.{part}"] = rot[idx].clone()
    return state


def save_checkpoint(run, path, step: int | None = None, sim_time: float | None = None) -> None:
    """Write the restart file of `run` (reference ``main.save_checkpoint``)."""
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
    dirk = getattr(run, "dirk", False)
    cond_ic = getattr(run, "cond_ic", False)
    rot_ic = getattr(run, "rot_ic", False)
    nold, nexp, nimp = slot_counts(run)
    scheme = "BPR353" if dirk else "CNAB2"
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
    if cond_ic:
        shape_ic = tuple(run.fic["b_ic"].shape)
        for name in IC_FIELD_NAMES:
            v = state.get(f"fields.{name}")
            if v is None:   # the reference keeps the run's inner fields then (main.py:239-242)
                continue
            if tuple(v.shape) != shape_ic:
                raise ValueError(f"load_checkpoint: fields.{name} has shape {tuple(v.shape)}, "
                                 f"run has {shape_ic}")
            run.fic[name].copy_(v.to(ref.device, ref.dtype))
    _check_zero(state, "fields.omega_ma")
    if not rot_ic:
        _check_zero(state, "fields.omega_ic")
        _check_zero(state, "dt.domega_ic_dt.impl")
        _check_zero(state, "dt.domega_ic_dt.expl")
        _check_zero(state, "dt.domega_ic_dt.old")
    for part in ("impl", "expl", "old"):
        _check_zero(state, f"dt.domega_ma_dt.{part}")

    def restore(name, old, impl, expl, shp):
        vi, ve, vo = (state[f"dt.{name}.{part}"] for part in ("impl", "expl", "old"))
        if (tuple(ve.shape) != shp + (nexp,) or tuple(vi.shape) != shp + (nimp,)
                or tuple(vo.shape) != shp + (nold,)):
            raise ValueError(f"load_checkpoint: dt.{name} arrays do not fit a {scheme} run of "
                             f"shape {shp}")
        old.copy_(vo[..., 0].to(ref.device, ref.dtype))
        for q, t in enumerate(impl):
            t.copy_(vi[..., q].to(ref.device, ref.dtype))
        # CNAB2: the next step overwrites the newest slot and reads slot 1 as its history
        for q, t in enumerate(expl):
            t.copy_(ve[..., q].to(ref.device, ref.dtype))

    for name, key in DT_NAMES.items():
        if key is None or (key in ("b", "j") and not run.p.l_mag):
            for part in ("impl", "expl", "old"):
                _check_zero(state, f"dt.{name}.{part}")
            continue
        restore(name, *_slots(run, key), shape)
    if cond_ic:
        for name, key in IC_DT_NAMES.items():
            if f"dt.{name}.old" in state:
                restore(name, *_slots(run, key, inner=True), shape_ic)
        if rot_ic:
            rot = torch.zeros(16, dtype=torch.float64)
            rot[0] = float(state["fields.omega_ic"])
            for part, idx, n in zip(("old", "impl", "expl"), _rot_slots(run), (nold, nimp, nexp)):
                v = state.get(f"dt.domega_ic_dt.{part}")
                if v is None:
                    continue
                if tuple(v.shape) != (n,):
                    raise ValueError(f"load_checkpoint: dt.domega_ic_dt.{part} does not fit a "
                                     f"{scheme} run")
                rot[idx] = v.to(torch.float64)
            run.rot.copy_(rot.to(run.rot.device))
    dt = state["tscheme_dt"]
    if dirk:
        run.dt_hist = [float(dt[0]), float(dt[0])]
    else:
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
    # (BPR353: the tableau's diagonal dt/2)
    run.build_all_matrices(0.5 * run.dt_hist[0] if dirk else run.p.alpha * run.dt_hist[0])
    return run.n_step, run.time


def load_checkpoint(run, path) -> tuple[int, float]:
    """Restore `run` from a restart file (reference ``main.load_checkpoint``);
    returns (step, sim_time).  Continue with ``run.one_step(step + 1, dt)``,
    dt = ``run.dt_hist[0]``."""
    return apply_checkpoint_state(run, torch.load(path, map_location="cpu", weights_only=True))
