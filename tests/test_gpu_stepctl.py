"""Device step control (``mh_step_control``) against the host decisions.

The kernel takes the decisions the host takes between the radial and the lm
phase of a CNAB2 step (Courant limits, the hysteresis rule of dt_courant,
``courant.py:80-155``; the dt roll and rebuild rule of one_step,
``step_time.py:1574-1637``; the CNAB2 weights, ``time_scheme.py:56-109``).
Host and device must agree bit for bit, so that a device-controlled run takes
the same steps as a host-controlled one.
"""

import ctypes as C
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _host_decision(v, delxr2, delxh2, st, alpha, BuoFac, sched=None):
    """The host path of DynamoRun (_cfl_values, dt_courant, finish_step, lm_loop)."""
    from magic_amd.dynamo import dt_courant
    N = delxr2.size
    with np.errstate(divide="ignore"):
        dtrkc = np.where(v[:N] > 0, np.sqrt(delxr2 / np.where(v[:N] > 0, v[:N], 1.0)), 1e10)
        dthkc = np.where(v[N:] > 0, np.sqrt(delxh2 / np.where(v[N:] > 0, v[N:], 1.0)), 1e10)
    n = st["n_step"] + 1
    new, dt_new = dt_courant(st["dt"], st["dtMax"], float(dtrkc.min()), float(dthkc.min()))
    if sched is not None and n <= len(sched):
        dt_new = float(sched[n - 1])
        new = dt_new != st["dt"]
    dt1, dt2 = dt_new, st["dt1"]
    wl0 = alpha * dt_new
    rebuild = n == 1 or new or math.isnan(st["wimp_lin0"])
    w = [1.0, (1.0 - alpha) * dt1, (1.0 + 0.5 * dt1 / dt2) * dt1, -0.5 * dt1 * dt1 / dt2,
         alpha * dt1 * BuoFac]
    return dict(dt=dt_new, new_dt=int(new), rebuild=int(rebuild), dt1=dt1, dt2=dt2,
                wimp_lin0=wl0 if rebuild else st["wimp_lin0"], w=w, n_step=n)


def _device_decision(v, delxr2, delxh2, st, alpha, BuoFac, sched=None, info=None, log=None):
    import torch
    from magic_amd import _lib
    from magic_amd.device import ptr, stream_handle
    dev = "cuda"
    N = delxr2.size
    s = _lib.StepState()
    for k in ("dt", "dtMax", "dt1", "dt2", "wimp_lin0", "time"):
        setattr(s, k, st.get(k, 0.0))
    for k in ("n_step", "rebuild", "n_rebuild", "sing_step"):
        setattr(s, k, st.get(k, 0))
    sd = torch.frombuffer(bytearray(bytes(s)), dtype=torch.uint8).to(dev)
    cfl = torch.as_tensor(v.view(np.int64).copy()).to(dev)
    dr, dh = torch.as_tensor(delxr2).to(dev), torch.as_tensor(delxh2).to(dev)
    w = torch.zeros(8, dtype=torch.float64, device=dev)
    a = _lib.StepCtlArgs()
    a.cfl, a.delxr2, a.delxh2, a.state, a.w = ptr(cfl), ptr(dr), ptr(dh), ptr(sd), ptr(w)
    if sched is not None:
        sch = torch.as_tensor(np.asarray(sched, np.float64)).to(dev)
        a.dt_schedule, a.n_schedule = ptr(sch), len(sched)
    if log is not None:
        a.dt_log, a.log_cap = ptr(log), log.numel()
    if info is not None:
        a.info, a.info_rows, a.info_cols = ptr(info), info.shape[0], info.shape[1]
    a.N, a.alpha, a.BuoFac = N, alpha, BuoFac
    _lib.call("mh_step_control", _lib.ref(a), stream_handle())
    torch.cuda.synchronize()
    out = _lib.StepState.from_buffer_copy(sd.cpu().numpy().tobytes())
    return out, w.cpu().numpy()


def _check(v, delxr2, delxh2, st, sched=None, alpha=0.6, BuoFac=3.7):
    h = _host_decision(v, delxr2, delxh2, st, alpha, BuoFac, sched)
    d, w = _device_decision(v, delxr2, delxh2, st, alpha, BuoFac, sched)
    assert d.dt == h["dt"] and d.dt1 == h["dt1"] and d.dt2 == h["dt2"], (d.dt, h["dt"])
    assert d.new_dt == h["new_dt"] and d.rebuild == h["rebuild"] and d.n_step == h["n_step"]
    assert d.wimp_lin0 == h["wimp_lin0"] or (math.isnan(d.wimp_lin0) and math.isnan(h["wimp_lin0"]))
    assert list(w[:5]) == h["w"], (list(w[:5]), h["w"])
    assert d.time == st.get("time", 0.0) + h["dt"]
    return d


def test_step_control_matches_host_on_every_branch():
    rng = np.random.default_rng(5)
    for N in (9, 33, 97, 300):
        delxr2 = rng.uniform(1e-4, 1e-2, N) ** 2
        delxh2 = rng.uniform(1e-3, 1e-1, N) ** 2
        v = rng.uniform(0.0, 1e4, 2 * N)
        v[rng.integers(0, 2 * N, N // 3)] = 0.0          # levels at rest
        with np.errstate(divide="ignore"):
            lim = min(np.sqrt(delxr2 / np.where(v[:N] > 0, v[:N], np.inf)).min(),
                      np.sqrt(delxh2 / np.where(v[N:] > 0, v[N:], np.inf)).min())
        base = dict(dtMax=1e-3, dt1=2e-5, dt2=2.5e-5, wimp_lin0=0.6 * 2e-5, n_step=7)
        for dt in (2e-3, 1.01 * lim, 0.6 * lim, 0.3 * lim, 0.45 * lim):  # > dtMax, > dtMin, hold, grow
            _check(v, delxr2, delxh2, dict(base, dt=dt))
        _check(v, delxr2, delxh2, dict(base, dt=0.6 * lim, n_step=0, wimp_lin0=math.nan))
        _check(np.zeros(2 * N), delxr2, delxh2, dict(base, dt=2e-5))   # no flow: 1e10 limits
        # replayed step sizes (inside and past the schedule)
        _check(v, delxr2, delxh2, dict(base, dt=2e-5), sched=[1e-5] * 7 + [2e-5, 3e-5])
        _check(v, delxr2, delxh2, dict(base, dt=2e-5, n_step=9), sched=[1e-5] * 7 + [2e-5, 3e-5])


def test_step_control_log_ring_and_singular_scan():
    import torch
    N = 33
    rng = np.random.default_rng(9)
    delxr2, delxh2 = rng.uniform(1e-4, 1e-2, N) ** 2, rng.uniform(1e-3, 1e-1, N) ** 2
    v = rng.uniform(1.0, 1e4, 2 * N)
    log = torch.full((4,), -1.0, dtype=torch.float64, device="cuda")
    info = torch.zeros(6, 17, dtype=torch.int32, device="cuda")
    st = dict(dt=1e-5, dtMax=1e-3, dt1=1e-5, dt2=1e-5, wimp_lin0=0.6e-5, n_step=5, rebuild=1)
    # clean pivots: nothing recorded; dt of step 6 lands at (6 - 1) % 4
    d, _ = _device_decision(v, delxr2, delxh2, st, 0.6, 1.0, info=info, log=log)
    assert d.sing_step == 0 and log.cpu().numpy()[1] == d.dt
    # two singular systems: the first in (system, degree) order is kept
    info[4, 3] = 12
    info[1, 9] = 2
    d, _ = _device_decision(v, delxr2, delxh2, st, 0.6, 1.0, info=info)
    assert (d.sing_step, d.sing_sys, d.sing_l, d.sing_info) == (5, 1, 9, 2)
    # the previous step did not rebuild: its flags are stale and not scanned
    d, _ = _device_decision(v, delxr2, delxh2, dict(st, rebuild=0), 0.6, 1.0, info=info)
    assert d.sing_step == 0
