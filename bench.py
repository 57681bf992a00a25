#!/usr/bin/env python
"""Benchmark: steps/s of the Boussinesq CNAB2 dynamo step on one MI355X.

Workload (the headline configuration): l_max=383, n_r_max=97 Chebyshev,
n_theta=576, n_phi=1152, E=1e-5, Ra=2e8 (modified Rayleigh 2000 / E), Pr=1,
Pm=1, radius ratio 0.35, CNAB2 with adaptive Courant dt (dt <= 0.15 E),
seeded Gaussian noise 1e-4 on the entropy plus the dipole seed field.
Synthetic data: fields initialised on the device from the seed, no I/O.

A "step" is one full ``one_step`` (radial loop with the SHTs, nonlinear
terms and CFL, dt decision, operator rebuild if dt changed, implicit solves).
W warm-up steps are run untimed, then exactly K steps are timed between two
device synchronisations.  Per-section HIP-event timings from the same timed
region give the roofline of the dominant kernel.  The CPU baseline is the
FP64 numpy oracle on a bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "torch-magic_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

WORKLOAD = dict(l_max=383, n_r_max=97, ek=1e-5, ra=2e8, pr=1.0, prmag=1.0, radratio=0.35,
                dtmax=1e-4)
NOISE = dict(seed=20261202, amp=1e-4)

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def section_work(p, N):
    """Algorithmic FLOPs (MFMA-bound sections) or HBM bytes (bandwidth-bound) per launch."""
    L, H, T, P, M = p.lm_max, p.n_theta // 2, p.n_theta, p.n_phi, p.n_m
    Nb = N - 2
    return {
        # Legendre: 4 flops per real-matrix x complex-coefficient multiply-add
        # with the equatorial fold (one H-length product per (l, m))
        "leg_syn_vec": ("mfma", 20.0 * L * H * 4 * N),
        "leg_syn_scal": ("mfma", 4.0 * L * H * N),
        "leg_ana_scal": ("mfma", 4.0 * L * H * 3 * Nb),
        "leg_ana_vec": ("mfma", 16.0 * L * H * 3 * Nb),
        # per-degree fused solves: (3N x N) on 4 systems, (6N x 2N) on (w, p)
        "rowop_szbj": ("mfma", 4 * 12.0 * L * N * N),
        "rowop_wp": ("mfma", 48.0 * L * N * N),
        # FFTs and grid products: bytes read + written once
        "fft_inv": ("hbm", 13.0 * N * T * (M * 16.0 + P * 8.0)),
        "fft_fwd": ("hbm", 9.0 * Nb * T * (P * 8.0 + M * 16.0)),
        "get_nl": ("hbm", (13.0 * N + 9.0 * Nb) * T * P * 8.0),
        # fused grid step: Fourier rows in (13 components, 6 on the 2 boundary
        # levels) and out (9 components, bulk levels), M complex values each
        "grid_step": ("hbm", ((13.0 * Nb + 6.0 * 2) + 9.0 * Nb) * T * M * 16.0),
    }


def cpu_baseline(sample_levels):
    """Time the FP64 oracle on a bounded sample of the workload (see 'sample')."""
    import numpy as np
    from oracle.dynamo import OracleRun
    cfg = dict(WORKLOAD, n_r_max=sample_levels)
    run = OracleRun(cfg)
    run.initialize_fields()
    run.add_noise(NOISE["seed"], NOISE["amp"])
    run.setup_initial_state()
    run.initialize_dt(WORKLOAD["dtmax"])
    run.one_step(1, WORKLOAD["dtmax"])   # builds the matrices
    t0 = time.perf_counter()
    run.one_step(2, WORKLOAD["dtmax"])
    dt = time.perf_counter() - t0
    try:
        import torch
        cores = torch.get_num_threads()
    except Exception:
        cores = os.cpu_count()
    return dict(value=1.0 / dt, unit="steps/s", cores=cores, kind="port",
                sample=f"one CNAB2 step of the numpy FP64 oracle at l_max=383 with n_r_max="
                       f"{sample_levels} (the full 576x1152 angular grid); the GPU value "
                       f"is at n_r_max=97, so scale by ~{sample_levels}/97 for equal work",
                seconds=dt)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-levels", type=int, default=9)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from magic_amd.dynamo import DynamoRun

    run = DynamoRun(WORKLOAD, device="cuda:0")
    run.initialize_fields(noise_amp=NOISE["amp"], noise_seed=NOISE["seed"])
    run.setup_initial_state()
    run.initialize_dt(WORKLOAD["dtmax"])
    dt = run.dtMax
    n = 0
    for _ in range(args.warmup):
        n += 1
        dt = run.one_step(n, dt)
    torch.cuda.synchronize()
    run.enable_timers(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        n += 1
        dt = run.one_step(n, dt)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    secs = run.timer_summary()
    run.enable_timers(False)
    elapsed = t1 - t0
    ms_per_step = 1e3 * elapsed / args.steps
    energies = run.energies()
    p, N = run.p, run.N
    work = section_work(p, N)
    # dominant kernel section by total time in the timed region
    dom = max(secs, key=lambda k: secs[k][0] * secs[k][1])
    cnt, ms = secs[dom]
    kind, amount = work.get(dom, ("hbm", float("nan")))
    if kind == "mfma":
        ach = amount / (ms * 1e-3) / 1e12
        roof = dict(bound="mfma", achieved=ach, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                    frac=ach / PEAK_FP64_TFLOPS, traffic=None, kernel=dom,
                    launch_ms=ms, algorithmic_flops=amount)
    else:
        ach = amount / (ms * 1e-3) / 1e9
        roof = dict(bound="hbm", achieved=ach, peak=PEAK_HBM_GBS, unit="GB/s",
                    frac=ach / PEAK_HBM_GBS, traffic=None, kernel=dom,
                    launch_ms=ms, algorithmic_bytes=amount)
    breakdown = {k: round(v[1], 4) for k, v in sorted(secs.items(), key=lambda kv: -kv[1][1])}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_levels)
    line = {
        "metric": "steps/s at l_max=383 mixed vs FP64; ensemble p50/p99 step latency",
        "value": args.steps / elapsed,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded noise + dipole seed, generated on device)",
        "config": {"workload": "Boussinesq CNAB2 dynamo step, l_max=383, n_r_max=97, "
                               "n_theta=576, n_phi=1152, E=1e-5, Ra=2e8, Pr=1, Pm=1, FP64",
                   "l_max": 383, "n_r_max": 97, "precision": "fp64", "dt": dt,
                   "rebuilds": run.n_rebuild, "energies": energies,
                   "section_ms": breakdown},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
