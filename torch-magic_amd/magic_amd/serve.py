"""Ensemble serving: many independent dynamo runs resident on one GPU.

A request is one dynamo run (configuration, initial-condition seed, number of
steps) arriving at a given time.  The server

* admits requests in arrival order while a slot is free and the run's HBM
  footprint fits the budget (admission by footprint, continuous: a request
  that arrives while others are stepping is admitted as soon as a slot and
  memory free up);
* keeps every admitted run's spectral state, IMEX history and operators
  resident in HBM until its last step, each run on its own HIP stream (default
  priority), with the step captured as HIP graphs;
* steps the active runs in rounds (every run advances one step per round, as
  iteration-level batching): a step is the radial phase (queued with the
  asynchronous copy of its Courant maxima) followed, once those maxima are on
  the host, by the dt decision and the lm phase; all runs' radial phases are
  queued before any host wait, so the runs' kernels overlap on the device;
* reports runs/hour and the p50/p99 per-step latency (wall time between the
  ends of a run's consecutive rounds).

The reference runs one configuration per process (``main.run``,
``main.py:280-993``); each run here follows the same step sequence
(``one_step``, ``step_time.py:1574-1637``), so its result does not depend on
what else is being served.  Runs sharing (l_max, n_r_max, radratio) share the
device Legendre tables.
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np


@dataclass
class Request:
    rid: int
    cfg: dict
    n_steps: int
    arrival: float = 0.0                 # seconds after the server starts
    noise_amp: float = 1e-5
    noise_seed: int = 0


@dataclass
class Result:
    rid: int
    cfg: dict
    n_steps: int
    arrival: float
    admitted: float
    finished: float
    step_latencies: list = field(default_factory=list)
    energies: tuple = ()
    footprint: int = 0
    final_fields: dict | None = None


def config5_requests(n_runs=64, rate=2.0, n_steps=500, seed=20261202):
    """The ensemble workload: Poisson arrivals at `rate` per second; l_max from
    {32, 64, 85, 127}, n_r from {33, 49, 65}, E log-uniform in [1e-4, 1e-3],
    modified Rayleigh number uniform in [60, 200] (ra = Ra_mod / E); Boussinesq
    dynamo (Pr=1, Pm=5, radius ratio 0.35, CNAB2, dtmax 1e-4), seeded noise."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n_runs):
        t += rng.exponential(1.0 / rate)
        l_max = int(rng.choice([32, 64, 85, 127]))
        n_r = int(rng.choice([33, 49, 65]))
        ek = float(10.0 ** rng.uniform(-4.0, -3.0))
        ra_mod = float(rng.uniform(60.0, 200.0))
        cfg = dict(l_max=l_max, n_r_max=n_r, ek=ek, ra=ra_mod / ek, pr=1.0, prmag=5.0,
                   radratio=0.35, dtmax=1e-4)
        out.append(Request(i, cfg, n_steps, arrival=t, noise_amp=1e-5, noise_seed=1000 + i))
    return out


def footprint_bytes(cfg) -> int:
    """Estimated HBM bytes of one resident run (fields, IMEX history, right-hand
    sides, nonlinear terms, Fourier buffers, operators; shared tables excluded)."""
    from .params import Params
    p = cfg if isinstance(cfg, Params) else Params(dict(cfg))
    L, N, M, T = p.lm_max, p.n_r_max, p.n_m, p.n_theta
    spec = L * N * 16
    n_spec = 15 + 3 + 6 * 4 + 4 + 2 + 10          # fields, tmp, old/impl/expl x2, rhs, nl
    fourier = M * T * 16 * (N + 3 * 4 * N + 9 * (N - 2))
    K4, K4wp = -(-N // 4) * 4, -(-(2 * N) // 4) * 4
    Rpad3, Rpad6 = -(-(3 * N) // 64) * 64, -(-(6 * N) // 64) * 64
    ops = (p.l_max + 1) * 8 * (4 * K4 * Rpad3 + K4wp * Rpad6)
    solve_work = (p.l_max + 1) * 8 * (N * N + 4 * N * N) * 2
    return int(spec * n_spec + fourier + ops + solve_work + (16 << 20))


class _Slot:
    def __init__(self, req: Request, run, stream, t_adm: float, fp: int):
        self.req, self.run, self.stream = req, run, stream
        self.n = 0                      # completed steps
        self.dt = run.dtMax
        self.last_done = t_adm
        self.res = Result(req.rid, req.cfg, req.n_steps, req.arrival, t_adm, 0.0, footprint=fp)


class EnsembleServer:
    """Continuous-batching server for dynamo runs on one GPU.

    hbm_budget: bytes available to resident runs (default: 80% of the free
    device memory at start); max_slots: runs stepping at once; use_graphs:
    replay each run's phases as HIP graphs.
    """

    def __init__(self, device="cuda", hbm_budget: int | None = None, max_slots: int = 16,
                 use_graphs: bool = True, clock=time.perf_counter):
        import torch
        from .device import require_gpu
        self.torch = torch
        self.device = require_gpu(device)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.8 * free)
        self.budget = hbm_budget
        self.max_slots = max_slots
        self.use_graphs = use_graphs
        self.clock = clock
        self._res_cache = {}

    def _resolution(self, p):
        from .device import Resolution
        key = (p.l_max, p.n_r_max, p.radratio, p.n_cheb_max, p.opm)  # opm: Alfven Courant factors
        r = self._res_cache.get(key)
        if r is None:
            r = self._res_cache[key] = Resolution(p, self.device)
        return r

    def _admit(self, req: Request, fp: int) -> _Slot:
        """Build and initialise a run on its own stream."""
        from .dynamo import DynamoRun
        from .params import Params
        torch = self.torch
        t0 = self.clock()
        p = Params(dict(req.cfg))
        stream = torch.cuda.Stream(device=self.device)
        with torch.cuda.stream(stream):
            run = DynamoRun(p, self.device, res=self._resolution(p))
            run.use_graphs = self.use_graphs
            run.defer_checks = True  # singular pivots surface at the next step, without a host wait
            run.initialize_fields(noise_amp=req.noise_amp, noise_seed=req.noise_seed)
            run.setup_initial_state()
            run.initialize_dt(p.dtmax)
        self.admit_s.append(self.clock() - t0)
        return _Slot(req, run, stream, 0.0, fp)

    def serve(self, requests, realtime: bool = True, keep_fields: bool = False):
        """Serve `requests` (arrival times honoured when realtime); returns the
        list of Result in completion order and a summary dict (keep_fields:
        attach copies of each run's final s, w, z, b, aj fields).

        Work proceeds in rounds, as iteration-level batching: in a round every
        active run advances exactly one step (all radial phases are queued
        first, then each run's lm phase as soon as its Courant maxima arrive),
        and arrivals are admitted between rounds.  A run's step latency is the
        time between the ends of its consecutive rounds (for its first step:
        from admission)."""
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        for r in pending:
            if footprint_bytes(r.cfg) > self.budget:
                raise MemoryError(f"request {r.rid} needs {footprint_bytes(r.cfg)} bytes, "
                                  f"more than the whole budget {self.budget}")
        active: list[_Slot] = []
        done: list[Result] = []
        used = 0
        self.rounds = []
        self.admit_s = []  # host time of each admission (run construction and set-up)
        t0 = self.clock()
        while pending or active:
            now = self.clock() - t0
            # admission between rounds: arrival order, by free slot and footprint
            while pending and len(active) < self.max_slots:
                req = pending[0]
                if realtime and req.arrival > now:
                    break
                fp = footprint_bytes(req.cfg)
                if used + fp > self.budget:
                    break
                pending.pop(0)
                used += fp
                sl = self._admit(req, fp)
                sl.last_done = sl.res.admitted = self.clock() - t0
                active.append(sl)
            if not active:
                time.sleep(max(0.0, min(0.01, pending[0].arrival - (self.clock() - t0))))
                continue
            used -= self._round(active, done, t0, keep_fields)
        wall = self.clock() - t0
        return done, summarize(done, wall)

    def _round(self, active, done, t0, keep_fields) -> int:
        """One round: every active run advances one step (all radial phases are
        queued first, then each run's lm phase as soon as its Courant maxima
        arrive).  Finished runs leave; returns the HBM footprint they free."""
        torch = self.torch
        for sl in active:
            with torch.cuda.stream(sl.stream):
                sl.run.begin_step()
        waiting = list(active)
        while waiting:
            for sl in list(waiting):
                if sl.run.cfl_ready():
                    with torch.cuda.stream(sl.stream):
                        try:
                            sl.dt = sl.run.finish_step(sl.n + 1, sl.dt)
                        except AssertionError as e:
                            raise AssertionError(f"request {sl.req.rid} step {sl.n + 1} "
                                                 f"(dt {sl.dt:.6g}): {e}") from e
                    sl.n += 1
                    waiting.remove(sl)
            if waiting:
                time.sleep(0)
        t_end = self.clock() - t0
        self.rounds.append((t_end, len(active)))
        freed = 0
        for sl in list(active):
            sl.res.step_latencies.append(t_end - sl.last_done)
            sl.last_done = t_end
            if sl.n >= sl.req.n_steps:
                with torch.cuda.stream(sl.stream):
                    try:
                        sl.run._check_info()  # the last rebuild's pivots
                    except AssertionError as e:
                        raise AssertionError(f"request {sl.req.rid} step {sl.n}: {e}") from e
                    sl.res.energies = sl.run.energies()
                    if keep_fields:
                        sl.res.final_fields = {k: sl.run.f[k].clone()
                                               for k in ("s", "w", "z", "b", "aj")}
                sl.res.finished = self.clock() - t0
                done.append(sl.res)
                active.remove(sl)
                freed += sl.res.footprint
                sl.run = None
        return freed


def summarize(results, wall: float) -> dict:
    """Throughput and step-latency percentiles over all runs, and the latency
    percentiles of each resolution class (l_max, n_r_max) in the mix."""
    lat = np.concatenate([np.asarray(r.step_latencies, float) for r in results]) if results else \
        np.zeros(0)
    steps = int(lat.size)
    by_class = {}
    cls = lambda r: (r.cfg.get("l_max"), r.cfg.get("n_r_max"))
    for key in sorted({cls(r) for r in results}, key=str):
        x = np.concatenate([np.asarray(r.step_latencies, float) for r in results if cls(r) == key])
        if x.size:
            p50, p99 = np.percentile(x, 50), np.percentile(x, 99)
            by_class[f"l{key[0]}_n{key[1]}"] = dict(steps=int(x.size), p50_ms=float(p50 * 1e3),
                                                     p99_ms=float(p99 * 1e3),
                                                     p99_over_p50=float(p99 / p50))
    return dict(runs=len(results), steps=steps, wall_s=wall, by_class=by_class,
                runs_per_hour=3600.0 * len(results) / wall if wall > 0 else math.nan,
                steps_per_s=steps / wall if wall > 0 else math.nan,
                p50_ms=float(np.percentile(lat, 50) * 1e3) if steps else math.nan,
                p99_ms=float(np.percentile(lat, 99) * 1e3) if steps else math.nan,
                p99_over_p50=float(np.percentile(lat, 99) / np.percentile(lat, 50)) if steps else math.nan)


def serve_sequential(requests, device="cuda", use_graphs=True, keep_fields=False):
    """Baseline: the same runs one after another (no overlap, arrivals ignored)."""
    srv = EnsembleServer(device=device, max_slots=1, use_graphs=use_graphs)
    return srv.serve(requests, realtime=False, keep_fields=keep_fields)
