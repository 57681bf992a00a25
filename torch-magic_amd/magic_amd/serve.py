"""Ensemble serving: many independent dynamo runs resident on one GPU.

A request is one dynamo run (configuration, initial-condition seed, number of
steps) arriving at a given time.  The server

* admits requests in arrival order while a slot is free and the run's HBM
  footprint fits the budget (admission by footprint, continuous: a request
  that arrives while others are stepping is admitted as soon as a slot and
  memory free up).  Admission is done in stages (device buffers, operator
  storage, job tables, initial state, capture of the two step graphs) while
  the device executes a round of the running runs, so the host work of an
  admission overlaps their steps;
* keeps every admitted run's spectral state, IMEX history and operators
  resident in HBM until its last step, each run on its own HIP stream (default
  priority);
* steps the runs with every decision on the device: a step is one HIP graph
  (radial loop, Courant dt decision, operator rebuild when dt changes, lm
  loop), so the host never waits inside a step.  The active runs advance in
  rounds, one step each per round (iteration-level batching); their streams
  overlap on the device;
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
    solve_work = (p.l_max + 1) * 8 * (4 * N * N + 4 * N * N + (8 * N * N if 2 * N > 208 else 0))
    return int(spec * n_spec + fourier + ops + solve_work + (16 << 20))


class _Slot:
    def __init__(self, req: Request, fp: int):
        self.req, self.fp = req, fp
        self.run = self.stream = None
        self.n = 0                      # completed steps
        self.last_done = 0.0
        self.res = None


class EnsembleServer:
    """Continuous-batching server for dynamo runs on one GPU.

    hbm_budget: bytes available to resident runs (default: 80% of the free
    device memory at start); max_slots: runs stepping at once; use_graphs:
    replay each step as a HIP graph.
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

    def warm(self, cfgs):
        """Build the shared device tables of every resolution in `cfgs` and tune
        their grid-step launch shape ahead of serving (as a server loads what it
        serves before taking requests)."""
        from .dynamo import DynamoRun
        from .params import Params
        seen = set()
        for cfg in cfgs:
            p = Params(dict(cfg))
            res = self._resolution(p)
            key = (id(res), p.precision, p.l_mag)
            if key not in seen:  # a throw-away run times every launch shape once
                seen.add(key)
                DynamoRun(p, self.device, res=res).close()
        self.torch.cuda.synchronize(self.device)

    def _admit_stages(self, sl: _Slot):
        """Admission of one run in stages (a generator: one stage per next()):
        device buffers, operator storage, job tables (each on the run's own
        stream), initial state and device step control, then the capture of
        each of its two step graphs."""
        from .dynamo import DynamoRun
        from .params import Params
        torch = self.torch
        req = sl.req
        p = Params(dict(req.cfg))
        sl.stream = torch.cuda.Stream(device=self.device)
        res = self._resolution(p)
        run, build = DynamoRun.staged(p, self.device, res=res)
        while True:
            with torch.cuda.stream(sl.stream):
                done = next(build, StopIteration) is StopIteration
            if done:
                break
            yield
        yield
        with torch.cuda.stream(sl.stream):
            run.use_graphs = self.use_graphs
            run.defer_checks = True
            run.initialize_fields(noise_amp=req.noise_amp, noise_seed=req.noise_seed)
            run.setup_initial_state()
            run.initialize_dt(p.dtmax)
            run.ctl_begin(run.dtMax)
        if self.use_graphs:
            for par in (0, 1):
                yield
                with torch.cuda.stream(sl.stream):
                    run.prepare_graph("step", par)
        sl.run = run

    def serve(self, requests, realtime: bool = True, keep_fields: bool = False):
        """Serve `requests` (arrival times honoured when realtime); returns the
        list of Result in completion order and a summary dict (keep_fields:
        attach copies of each run's final s, w, z, b, aj fields).

        Work proceeds in rounds (iteration-level batching): every active run
        queues exactly one step (one HIP graph) on its stream; while the round
        runs on the device the host advances the admission in progress by
        stages (arrival order, by free slot and footprint); the round ends when
        every run's step event has completed.  A run's step latency is the
        time between the ends of its consecutive rounds (for its first step:
        from the end of its admission).  A run admitted during a round steps
        from the next round on."""
        torch = self.torch
        pending = sorted(requests, key=lambda r: (r.arrival, r.rid))
        for r in pending:
            if footprint_bytes(r.cfg) > self.budget:
                raise MemoryError(f"request {r.rid} needs {footprint_bytes(r.cfg)} bytes, "
                                  f"more than the whole budget {self.budget}")
        active: list[_Slot] = []
        done: list[Result] = []
        used = [0]
        admitting = [None]            # (slot, stage generator, start time)
        self.occupancy = []           # (time, active runs) at the end of each round
        self.admit_s = []             # host time of each admission (all its stages)
        self.stage_s = []             # host time of each admission stage
        t0 = self.clock()

        def admission_stage():
            """Start an admission if one may start, else advance the current one
            by a stage; True if host work was done."""
            if admitting[0] is None:
                if not pending or len(active) >= self.max_slots:
                    return False
                req = pending[0]
                fp = footprint_bytes(req.cfg)
                if (realtime and req.arrival > self.clock() - t0) or used[0] + fp > self.budget:
                    return False
                pending.pop(0)
                used[0] += fp
                sl = _Slot(req, fp)
                admitting[0] = (sl, self._admit_stages(sl), self.clock())
            sl, gen, ta = admitting[0]
            ts = self.clock()
            try:
                next(gen)
                self.stage_s.append(self.clock() - ts)
            except StopIteration:
                self.stage_s.append(self.clock() - ts)
                self.admit_s.append(self.clock() - ta)
                admitting[0] = None
                sl.last_done = self.clock() - t0
                sl.res = Result(sl.req.rid, sl.req.cfg, sl.req.n_steps, sl.req.arrival,
                                sl.last_done, 0.0, footprint=sl.fp)
                active.append(sl)
            return True

        while pending or active or admitting[0] is not None:
            stepping = list(active)
            if not stepping:
                if not admission_stage() and pending:
                    time.sleep(max(0.0, min(0.01, pending[0].arrival - (self.clock() - t0))))
                continue
            events = []
            for sl in stepping:
                with torch.cuda.stream(sl.stream):
                    try:
                        sl.run.step_async()
                    except AssertionError as e:  # a singular rebuild read back early
                        raise AssertionError(f"request {sl.req.rid}: {e}") from e
                    ev = torch.cuda.Event()
                    ev.record()
                events.append(ev)
            # host admission work overlaps the round on the device
            while not all(ev.query() for ev in events):
                if not admission_stage():
                    time.sleep(0)
            t_end = self.clock() - t0
            self.occupancy.append((t_end, len(stepping)))
            for sl in stepping:
                sl.res.step_latencies.append(t_end - sl.last_done)
                sl.last_done = t_end
                sl.n += 1
                if sl.n >= sl.req.n_steps:
                    self._retire(sl, keep_fields)
                    sl.res.finished = self.clock() - t0
                    done.append(sl.res)
                    active.remove(sl)
                    used[0] -= sl.fp
        wall = self.clock() - t0
        return done, summarize(done, wall)

    def _retire(self, sl: _Slot, keep_fields: bool):
        """Bring a finished run's state back (pivot checks), read its energies."""
        with self.torch.cuda.stream(sl.stream):
            try:
                sl.run.ctl_end()
            except AssertionError as e:
                raise AssertionError(f"request {sl.req.rid}: {e}") from e
            sl.res.energies = sl.run.energies()
            if keep_fields:
                sl.res.final_fields = {k: sl.run.f[k].clone() for k in ("s", "w", "z", "b", "aj")}
            sl.run.close()  # graphs and buffers released here, outside any graph capture
        sl.run = None


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


def serve_sequential(requests, device="cuda", use_graphs=True, keep_fields=False, warm=False):
    """Baseline: the same runs one after another (no overlap, arrivals ignored;
    warm: build the device tables of their resolutions first)."""
    srv = EnsembleServer(device=device, max_slots=1, use_graphs=use_graphs)
    if warm:
        srv.warm([r.cfg for r in requests])
    return srv.serve(requests, realtime=False, keep_fields=keep_fields)
