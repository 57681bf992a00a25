"""Launch-shape tuner.

The fused grid step (``mh_grid_step``) can give a workgroup 1, 2, ... up to
the most radial levels its FFT plan fits (``mh_grid_plan``).  The best count
depends on n_phi, the number of levels and the precision (registers, LDS per
workgroup, workgroups per theta row), and every count gives the same result,
so it is chosen by timing the candidates on the device, once per (device,
n_phi, n_theta, orders, levels, precision), and remembered: in memory for the
process and in a JSON file (``MAGIC_AMD_TUNE_CACHE``, default
``~/.cache/magic_amd/tune.json``; not written when the directory is not
writable).  ``MAGIC_AMD_TUNE=0`` keeps the plan's own level block.
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


def choose(times: dict[int, float], own_rb: int) -> int:
    """The fastest candidate as the `rb` argument (0 when it is the plan's own);
    a tie keeps the plan's own block."""
    best = min(times, key=lambda rb: (times[rb], rb != own_rb))
    return 0 if best == own_rb else best


def grid_key(res, N: int, fprec: int, l_mag: bool) -> str:
    name = torch.cuda.get_device_name(res.device).replace("|", "/")
    return f"{name}|grid_step|Mh={res.Mh}|T={res.T}|M={res.M}|N={N}|fprec={fprec}|mag={int(l_mag)}"


def grid_level_block(res, N: int, p, fdtype=torch.float64, bufs=None, reps: int = 3) -> int:
    """The `rb` to pass to ``Resolution.grid_step`` for N levels and Fourier
    buffers of `fdtype` (float64, or float32 in the mixed-precision step).

    bufs: (Fs, Fv, Fo, cfl) to time on (a run's own buffers); by default
    temporary ones are allocated.  Timing synchronises the current stream."""
    if os.environ.get("MAGIC_AMD_TUNE", "1") == "0":
        return 0
    fprec = 1 if fdtype == torch.float32 else 0
    own = res.grid_plan()[2]
    key = grid_key(res, N, fprec, p.l_mag)
    with _LOCK:
        if key in _MEM:
            return _MEM[key]
        hit = _read_file().get(key)
        if hit is not None and hit.get("rb", -1) in [0] + grid_candidates(own):
            _MEM[key] = int(hit["rb"])
            return _MEM[key]
    cands = grid_candidates(own)
    if len(cands) == 1:
        rb = 0
    else:
        if bufs is None:
            Fs = res.fourier_buffer(1, N, fdtype).normal_(0.0, 1e-3)
            Fv = res.fourier_buffer(3, 4 * N, fdtype).normal_(0.0, 1e-3)
            Fo = res.fourier_buffer(9, N - 2, fdtype)
            cfl = torch.zeros(2 * N, dtype=torch.int64, device=res.device)
        else:
            Fs, Fv, Fo, cfl = bufs
        times = {}
        for c in cands:
            res.grid_step(Fs, Fv, Fo, N, cfl, p, rb=0 if c == own else c)  # warm-up
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record()
            for _ in range(reps):
                res.grid_step(Fs, Fv, Fo, N, cfl, p, rb=0 if c == own else c)
            ev[1].record()
            ev[1].synchronize()
            times[c] = ev[0].elapsed_time(ev[1]) / reps
        rb = choose(times, own)
        _write_file(key, dict(rb=rb, ms={str(k): round(v, 4) for k, v in times.items()}))
    with _LOCK:
        _MEM[key] = rb
    return rb
