"""Launch-shape tuner.

Two launch shapes are chosen by timing on the device, once per shape and
precision, because the best one depends on the sizes and every choice gives
the same results bit for bit:

* the fused grid step (``mh_grid_step``): radial levels per workgroup, 1, 2,
  ... up to the most its FFT plan fits (``mh_grid_plan``);
* the Legendre synthesis (``mh_leg_synth``): waves per workgroup, 4 or 6
  (16 theta pairs each), and for the FP32 vector synthesis one or two column
  tiles per workgroup; the vector and the scalar synthesis are tuned apart;
* the Legendre analysis (``mh_leg_analysis``): waves per workgroup, 6 or 4
  (16 degrees each);
* the row operator (``mh_rowop``): tiles of 32 rows x 64 or x 128 columns.

Choices are remembered in memory for the process and in a JSON file
(``MAGIC_AMD_TUNE_CACHE``, default ``~/.cache/magic_amd/tune.json``; not
written when the directory is not writable).  ``MAGIC_AMD_TUNE=0`` keeps the
library defaults.
"""

from __future__ import annotations

import json
import os
import threading

import torch

_MEM: dict = {}
_LOCK = threading.Lock()


def cache_path() -> str:
    return os.environ.get("MAGIC_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "magic_amd", "tune.json"))


def _read_file() -> dict:
    try:
        with open(cache_path()) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def _write_file(key: str, entry: dict) -> None:
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        data = _read_file()
        data[key] = entry
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home: keep the choice in memory only


def grid_candidates(own_rb: int) -> list[int]:
    """Level blocks to try: the plan's own and the powers of two below it."""
    c = {own_rb}
    r = 1
    while r < own_rb:
        c.add(r)
        r *= 2
    return sorted(c)


def fastest(times: dict, default):
    """The candidate with the least time; a tie keeps `default`."""
    return min(times, key=lambda c: (times[c], c != default))


def timed_choice(key: str, candidates, launch, default, reps: int = 3, trials: int = 3):
    """The fastest of `candidates` for `key` (cached): launch(c) issues one run
    of candidate c on the current stream.  Each candidate's time is the least
    of `trials` interleaved measurements of `reps` runs (a transient slowdown
    cannot decide the choice).  Ties keep `default`."""
    with _LOCK:
        if key in _MEM:
            return _MEM[key]
        hit = _read_file().get(key)
        choice = hit.get("choice") if hit is not None else None
        choice = tuple(choice) if isinstance(choice, list) else choice  # JSON lists
        if choice is not None and choice in candidates:
            _MEM[key] = choice
            return choice
    if len(candidates) == 1:
        best = candidates[0]
    else:
        for c in candidates:
            launch(c)  # warm-up (code object load, first-touch)
        times = {}
        for _ in range(trials):
            for c in candidates:
                ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
                ev[0].record()
                for _ in range(reps):
                    launch(c)
                ev[1].record()
                ev[1].synchronize()
                t = ev[0].elapsed_time(ev[1]) / reps
                times[c] = min(t, times.get(c, t))
        best = fastest(times, default)
        _write_file(key, dict(choice=best, ms={str(k): round(v, 4) for k, v in times.items()}))
    with _LOCK:
        _MEM[key] = best
    return best


def _device_name(res) -> str:
    """The device in a cache key: architecture and compute-unit count (stable
    whatever name the runtime reports)."""
    pr = torch.cuda.get_device_properties(res.device)
    arch = getattr(pr, "gcnArchName", "") or pr.name
    return f"{arch.split(':')[0]}x{pr.multi_processor_count}".replace("|", "/")


def synth_shape(res, N: int, prec: int, fdtype=torch.float64, bufs=None,
                vector: bool = True) -> tuple:
    """(waves, column tiles) per workgroup for the dynamo step's synthesis:
    vector (4 groups of N levels) or scalar (one field of N levels), FP64
    (prec 0) or FP32 tables (prec 1) writing Fourier buffers of `fdtype`.
    bufs: (a spectral (L, N) array, the F buffer)."""
    if os.environ.get("MAGIC_AMD_TUNE", "1") == "0":
        return (0, 0)
    fprec = 1 if fdtype == torch.float32 else 0
    ng = 4 if vector else 1
    key = (f"{_device_name(res)}|leg_synth_{'vec' if vector else 'scal'}|L={res.L}|H={res.H}"
           f"|M={res.M}|B={ng * N}|prec={prec}|fprec={fprec}")
    if prec >= 2:  # BF16 split: four-wave workgroups only
        cands = [(4, 1)] + ([(4, 2)] if vector else [])
    else:
        cands = [(4, 1), (6, 1)] + ([(4, 2), (6, 2)] if (vector and prec == 1) else [])
    buf = {}

    def launch(c):  # buffers made on first use (none on a cache hit)
        if not buf:
            if bufs is None:
                buf["x"] = torch.zeros(res.L, N, dtype=torch.complex128, device=res.device)
                buf["F"] = res.fourier_buffer(3 if vector else 1, ng * N, fdtype)
            else:
                buf["x"], buf["F"] = bufs
        x = buf["x"]
        groups = ([[(x, None, 1), (x, None, 0), (x, x, 2)] for _ in range(4)] if vector
                  else [[(x, None, 0)]])
        res.synth(groups, N, N, buf["F"], ng * N, vector=vector, prec=prec, waves=c[0], nct=c[1])

    return timed_choice(key, cands, launch, (4, 1))


def analysis_waves(res, N: int, prec: int, fdtype=torch.float64, bufs=None) -> int:
    """Waves per workgroup for the vector analysis of 3 pairs of N-2 levels (the
    dynamo step's), FP64 (prec 0) or FP32 tables (prec 1) reading Fourier
    buffers of `fdtype`.  bufs: (the F buffer of 9 components, an (L, N)
    output array whose boundary columns 0 and N-1 are left untouched)."""
    if os.environ.get("MAGIC_AMD_TUNE", "1") == "0":
        return 0
    fprec = 1 if fdtype == torch.float32 else 0
    Nb = N - 2
    key = (f"{_device_name(res)}|leg_analysis_vec|L={res.L}|H={res.H}|M={res.M}|B={Nb}"
           f"|prec={prec}|fprec={fprec}")
    buf = {}

    def launch(w):  # buffers made on first use (none on a cache hit)
        if not buf:
            if bufs is None:
                buf["F"] = res.fourier_buffer(9, Nb, fdtype).zero_()
                buf["y"] = torch.zeros(res.L, N, dtype=torch.complex128, device=res.device)
            else:
                buf["F"], buf["y"] = bufs
        y = buf["y"].data_ptr() + 16  # columns 1 .. N-2, as the step writes them
        res.analysis(buf["F"], Nb, Nb, [(3, 4), (5, 6), (7, 8)], [y, y, y], [y, None, y], N,
                     vector=True, prec=prec, waves=w)

    if prec >= 2:
        return 4   # BF16 split: four-wave workgroups only
    return timed_choice(key, [6, 4], launch, 6)


def rowop_wide(run) -> int:
    """1 when the wide row-operator tiles (32 x 128) beat the narrow ones
    (32 x 64) on the run's S, Z, B, J solves.  Runs the solves on the run's
    current (zero) state, so call it before the fields are set."""
    if os.environ.get("MAGIC_AMD_TUNE", "1") == "0":
        return 0
    r = run.res
    key = f"{_device_name(r)}|rowop|L={r.L}|N={run.N}|l_max={run.p.l_max}|mag={int(run.p.l_mag)}"
    return timed_choice(key, [0, 1], lambda w: run._rowop(run.jobs1, w), 0)


def grid_key(res, N: int, fprec: int, l_mag: bool) -> str:
    return f"{_device_name(res)}|grid_step|Mh={res.Mh}|T={res.T}|M={res.M}|N={N}|fprec={fprec}|mag={int(l_mag)}"


def grid_level_block(res, N: int, p, fdtype=torch.float64, bufs=None, reps: int = 3) -> int:
    """The `rb` to pass to ``Resolution.grid_step`` for N levels and Fourier
    buffers of `fdtype` (float64, or float32 in the mixed-precision step).

    bufs: (Fs, Fv, Fo, cfl) to time on (a run's own buffers); by default
    temporary ones are allocated.  Timing synchronises the current stream."""
    if os.environ.get("MAGIC_AMD_TUNE", "1") == "0":
        return 0
    fprec = 1 if fdtype == torch.float32 else 0
    own = res.grid_plan()[2]
    cands = grid_candidates(own)
    if len(cands) == 1:
        return 0
    buf = []

    def launch(c):  # buffers made on first use (none on a cache hit)
        if not buf:
            buf.extend(bufs if bufs is not None else (
                res.fourier_buffer(1, N, fdtype).normal_(0.0, 1e-3),
                res.fourier_buffer(3, 4 * N, fdtype).normal_(0.0, 1e-3),
                res.fourier_buffer(9, N - 2, fdtype),
                torch.zeros(2 * N, dtype=torch.int64, device=res.device)))
        Fs, Fv, Fo, cfl = buf
        res.grid_step(Fs, Fv, Fo, N, cfl, p, rb=0 if c == own else c)

    best = timed_choice(grid_key(res, N, fprec, p.l_mag), cands, launch, own, reps)
    return 0 if best == own else best
