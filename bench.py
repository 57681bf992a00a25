#!/usr/bin/env python
"""Benchmark: steps/s of the Boussinesq CNAB2 dynamo step on one MI355X.

Workload (the headline configuration): l_max=383, n_r_max=97 Chebyshev,
n_theta=576, n_phi=1152, E=1e-5, Ra=2e8 (modified Rayleigh 2000 / E), Pr=1,
Pm=1, radius ratio 0.35, CNAB2 with adaptive Courant dt (dt <= 0.15 E),
seeded Gaussian noise 1e-4 on the entropy plus the dipole seed field.
Synthetic data: fields initialised on the device from the seed, no I/O.

A "step" is one full CNAB2 step (radial loop with the SHTs, nonlinear terms
and CFL, dt decision, operator rebuild if dt changed, implicit solves), run
as one HIP graph replay with the dt decision taken on the device
(``DynamoRun.step_async``), so the host queues steps without waiting.
W warm-up steps are run untimed, then exactly K steps (default 1000) are
timed between two device synchronisations, first in plain FP64, then in mixed
precision (Legendre transforms on FP32 MFMA, FP32 Fourier buffers, azimuthal
FFTs and nonlinear products; solves, radial derivatives, IMEX terms and
diagnostics FP64), each from the same seeded initial state; the mixed run replays the FP64 run's step
sizes so that the two differ by arithmetic precision only.  The headline value
is the mixed-precision steps/s; the FP64 rate, their ratio and the relative
energy and spectrum differences after W+K steps are in ``config``.
Per-section HIP-event timings, from 20 further steps run eagerly with events
between the sections, give the roofline of the dominant kernel.  Also reported: the SHT round trip at l_max=383 on 291 fields,
the finite-difference banded solves (FD4, n_r=512, l_max=255) and the 64-run ensemble served with continuous admission
against a sequential sample.  The CPU baseline is the FP64 numpy oracle on a
bounded sample of the same workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
                       [--precision both|all|fp64|mixed|mixed_bf16|mixed_bf16x3]
                       [--mixed mixed|mixed_bf16|mixed_bf16x3]
                       [--no-cpu-baseline] [--no-fd] [--no-ensemble] [--no-sht] [--no-peaks]

The measured peaks (a device copy, FP64 and FP32 GEMMs) are reported beside the
published ones, and the roofline gives the dominant kernel's fraction of both.
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

# HBM bytes per launch (read + written) of the main kernels at this workload, from
# rocprofv3 FETCH_SIZE / WRITE_SIZE passes scaled by a 256 MiB copy (FETCH x 2,
# WRITE x 1); see profiles/hbm_traffic_l383_nr97_pmc.md.  Grid-step entries are
# keyed by the tuner's radial levels per workgroup (0: the plan's own block, four
# levels for the FP32 grid step at n_phi = 1152; the FP64 step chose one); other
# level blocks have no measured figure (traffic null).  The FP32 Legendre entries
# ("mixed") are from the same kernels profiled before the BF16 headline.
PMC_TRAFFIC_BYTES = {
    ("grid_step", "mixed", 0): 2.540e9 + 1.791e9,
    ("grid_step", "fp64", 1): 5.861e9 + 4.554e9,
    ("leg_syn_vec", "mixed_bf16x3"): 3.444e9 + 2.171e9,
    ("leg_syn_vec", "mixed"): 4.196e9 + 2.192e9,
    ("leg_syn_vec", "fp64"): 5.721e9 + 4.187e9,
    ("leg_ana_vec", "mixed_bf16x3"): 1.640e9 + 0.601e9,
    ("leg_ana_vec", "mixed"): 1.414e9 + 0.599e9,
    ("leg_ana_vec", "fp64"): 7.779e9 + 0.620e9,
}

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
PEAK_BF16_TFLOPS = 2500.0   # dense BF16 matrix rate

MIXED = ("mixed", "mixed_bf16", "mixed_bf16x3")
# headline: three bf16 planes (energies within 1e-6 of FP64 after 1005 steps,
# and the fastest variant that meets that bar; two planes do not)
HEAD_MIXED = "mixed_bf16x3"
LEG_SPLIT = {
    "mixed": "Legendre synthesis/analysis on FP32 MFMA with FP32 accumulation",
    "mixed_bf16": "Legendre synthesis/analysis on BF16 MFMA, tables and inputs split into two "
                  "bf16 planes (hi*hi + hi*lo + lo*hi), FP32 accumulation",
    "mixed_bf16x3": "Legendre synthesis/analysis on BF16 MFMA, tables and inputs split into three "
                    "bf16 planes (the six products of plane pairs i + j < 3), FP32 accumulation",
}
DTYPE = {"fp64": "f64", "mixed": "f32+f64", "mixed_bf16": "bf16x2+f32+f64",
         "mixed_bf16x3": "bf16x3+f32+f64"}


def section_work(p, N, fbytes=16.0):
    """Algorithmic FLOPs (MFMA-bound sections) or HBM bytes (bandwidth-bound) per launch;
    fbytes: bytes per complex value of the Fourier buffers (16 FP64, 8 in mixed precision)."""
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
        "grid_step": ("hbm", ((13.0 * Nb + 6.0 * 2) + 9.0 * Nb) * T * M * fbytes),
    }


def cpu_baseline(sample_levels, steps=1, second_levels=17):
    """Time the FP64 oracle on a bounded sample of the workload: the full
    l_max=383 angular grid (576 x 1152) with n_r_max = sample_levels and with
    n_r_max = second_levels radial levels.  The line through the two samples,
    t(N) = a + b N, is reported as an extrapolation to n_r_max = 97; it is a
    lower bound: the per-degree solves and radial derivatives grow as N^2 and
    the larger arrays leave the caches (one full step at n_r_max = 97, 8 threads,
    took 271 s where the line through 9 and 17 levels gives 148 s).  A full step
    takes minutes on the host, too long for the default run."""
    from oracle.dynamo import OracleRun

    def per_step(levels, n):
        cfg = dict(WORKLOAD, n_r_max=levels)
        run = OracleRun(cfg)
        run.initialize_fields()
        run.add_noise(NOISE["seed"], NOISE["amp"])
        run.setup_initial_state()
        run.initialize_dt(WORKLOAD["dtmax"])
        dt_step = run.one_step(1, WORKLOAD["dtmax"])   # builds the matrices
        t0 = time.perf_counter()
        for k in range(2, 2 + n):
            dt_step = run.one_step(k, dt_step)
        return (time.perf_counter() - t0) / n

    t1 = per_step(sample_levels, steps)
    t2 = per_step(second_levels, steps)
    slope = (t2 - t1) / (second_levels - sample_levels)
    t97 = t1 + slope * (WORKLOAD["n_r_max"] - sample_levels)
    try:
        import torch
        cores = torch.get_num_threads()
    except Exception:
        cores = os.cpu_count()
    return dict(value=1.0 / t1, unit="steps/s", cores=cores, kind="port",
                sample=f"{steps} CNAB2 steps of the numpy FP64 oracle at l_max=383 (the full "
                       f"576x1152 angular grid) with n_r_max={sample_levels}; value is that "
                       f"sample's steps/s",
                seconds=t1,
                seconds_at_second_sample={str(second_levels): t2},
                extrapolated_seconds_per_step_at_97=t97,
                extrapolated_steps_per_s_at_97=1.0 / t97,
                extrapolation=f"linear in n_r_max through n_r_max={sample_levels} and "
                              f"{second_levels} (same angular grid), evaluated at 97: a lower "
                              f"bound (superlinear in n_r_max)")


def copy_calibration(reps=3):
    """The device copies that calibrate rocprofv3's FETCH_SIZE and WRITE_SIZE
    counters: `reps` copies of a 256 MiB float64 tensor (one read and one write
    of 256 MiB each), profiled on their own (``--copy-calibration``)."""
    import torch
    n = (256 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": 8 * n, "bytes_written": 8 * n,
                                           "copies": reps}}))


def run_case(precision, steps, warmup, dt_schedule=None, section_steps=20):
    """Time `steps` CNAB2 steps (one HIP graph each, dt decided on the device)
    after `warmup` untimed ones, then time the kernel sections over
    `section_steps` further steps run eagerly with HIP events between sections.
    dt_schedule: step sizes to replay instead of the run's own Courant decisions."""
    import torch
    from magic_amd.dynamo import DynamoRun
    run = DynamoRun(dict(WORKLOAD, precision=precision), device="cuda:0")
    run.dt_schedule = dt_schedule
    run.use_graphs = True
    run.initialize_fields(noise_amp=NOISE["amp"], noise_seed=NOISE["seed"])
    run.setup_initial_state()
    run.initialize_dt(WORKLOAD["dtmax"])
    dt = run.run_steps(warmup, run.dtMax)           # includes the graph captures
    run.ctl_begin(dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run.step_async()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dt = run.ctl_end()
    out = dict(elapsed=t1 - t0, energies=run.energies(), spectra=run.spectra(), dt=dt,
               rebuilds=run.n_rebuild, p=run.p, N=run.N, grid_rb=run.grid_rb,
               syn_shape=(run.syn_vec_shape, run.syn_scal_shape), ana_waves=run.ana_waves,
               rowop_wide=run.rowop_wide)
    run.enable_timers(True)
    run.run_steps(section_steps, dt)
    out["secs"] = run.timer_summary()
    out["dt_log"] = list(run.dt_log)
    run.enable_timers(False)
    run.close()
    del run
    torch.cuda.empty_cache()
    return out


# The reference's own published FP64 step times on NVIDIA parts (dynamo
# benchmark physics, eager PyTorch on the GPU): l_max=64/n_r=129 Chebyshev
# (PROGRESS.md:2252) and finite differences FD2 (perf.md:249-257) on an H100,
# l_max=128/n_r=257 Chebyshev on a B200 (perf.md:270-283), and the operator
# rebuild on a dt change at l_max=64 on an H100 (PROGRESS.md:2254).
PUBLISHED = [
    dict(name="chebyshev_l64_n129", cfg=dict(l_max=64, n_r_max=129), ms=6.7, hw="H100"),
    dict(name="fd2_l64_n129", cfg=dict(l_max=64, n_r_max=129, radial_scheme="FD", fd_order=2,
                                       fd_order_bound=2), ms=9.9, hw="H100"),
    dict(name="chebyshev_l128_n257", cfg=dict(l_max=128, n_r_max=257), ms=56.0, hw="B200"),
]
PUBLISHED_REBUILD_MS = 70.0   # l_max=64, n_r=129, H100


def published_case(steps=100, warmup=10):
    """The reference's published step-time points, run here in FP64 with the
    same physics (E=1e-3, Ra=1e5, Pr=1, Pm=5, radius ratio 0.35, CNAB2, dt 1e-4,
    benchmark initial fields): ms per step over `steps` graph-replayed steps
    with the dt decision on the device, and (l_max=64 Chebyshev) the mean time
    of a full operator rebuild."""
    import torch
    from magic_amd.dynamo import DynamoRun
    phys = dict(ek=1e-3, ra=1e5, pr=1.0, prmag=5.0, radratio=0.35, dtmax=1e-4)
    out = {}
    for pt in PUBLISHED:
        run = DynamoRun(dict(phys, precision="fp64", **pt["cfg"]), device="cuda:0")
        run.use_graphs = True
        run.initialize_fields()
        run.setup_initial_state()
        run.initialize_dt(phys["dtmax"])
        dt = run.run_steps(warmup, run.dtMax)
        run.ctl_begin(dt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            run.step_async()
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        run.ctl_end()
        rec = dict(ms_per_step=ms, published_ms=pt["ms"], published_hw=pt["hw"],
                   speedup=pt["ms"] / ms, energies=run.energies())
        if pt["name"] == "chebyshev_l64_n129":
            w = run.p.alpha * run.dtMax     # CNAB2 implicit weight at the run's dt
            run.build_all_matrices(w)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(5):
                run.build_all_matrices(w)
            torch.cuda.synchronize()
            rb = 1e3 * (time.perf_counter() - t0) / 5
            rec.update(rebuild_ms=rb, published_rebuild_ms=PUBLISHED_REBUILD_MS,
                       rebuild_speedup=PUBLISHED_REBUILD_MS / rb)
        out[pt["name"]] = rec
        print(f"bench: published point {pt['name']}: {ms:.3f} ms/step "
              f"(reference {pt['ms']} ms on {pt['hw']})", file=sys.stderr, flush=True)
        run.close()
        del run
        torch.cuda.empty_cache()
    return out


def fd_step_case(steps=20, warmup=3):
    """Config 3's physics as a full time step: finite differences FD4 (boundary
    order 2) on the stretched n_r=512 grid, l_max=255, anelastic polytrope
    N_rho=3 (strat 3, polind 2, gravity ~ 1/r^2), hydrodynamic, stress-free
    walls, E=1e-3, Ra=1e5, Pr=1, FP64, CNAB2 with the double-curl poloidal
    equation: ms per step (host dt decision, eager launches) and the time of
    the first operator build (host assembly excluded)."""
    import torch
    from magic_amd.dynamo import DynamoRun
    cfg = dict(l_max=255, n_r_max=512, radial_scheme="FD", fd_order=4, fd_order_bound=2,
               strat=3.0, polind=2.0, g0=0.0, g1=0.0, g2=1.0, mode=1, ktopv=1, kbotv=1,
               ra=1e5, ek=1e-3, pr=1.0, dtmax=1e-5, precision="fp64")
    t0 = time.perf_counter()
    run = DynamoRun(cfg, device="cuda:0")
    setup_s = time.perf_counter() - t0
    run.initialize_fields(noise_amp=NOISE["amp"], noise_seed=NOISE["seed"])
    run.setup_initial_state()
    run.initialize_dt(cfg["dtmax"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run.build_all_matrices(run.p.alpha * run.dtMax)
    torch.cuda.synchronize()
    build_ms = 1e3 * (time.perf_counter() - t0)
    dt = run.run_steps(warmup, run.dtMax)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dt = run.run_steps(steps, dt)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    out = dict(ms_per_step=ms, steps=steps, operator_build_ms=build_ms, setup_s=setup_s,
               rebuilds=run.n_rebuild, energies=run.energies(), dt=dt,
               workload="FD4 n_r=512 l_max=255 anelastic N_rho=3 hydro, FP64")
    run.close()
    del run
    torch.cuda.empty_cache()
    return out


def fd_banded_case():
    """Config 3: FD4 n_r=512 l_max=255 anelastic (N_rho=3, polytropic index 2)
    banded S, Z, W operators: batched band LU of every degree and the solve of
    all 32 896 complex right-hand sides per system in one launch each."""
    import numpy as np
    import torch
    from magic_amd.fd import FDOperators
    n, l_max = 512, 255
    t0 = time.perf_counter()
    ops = FDOperators(n, l_max, 0.6 * 1e-4, fd_order=4, fd_order_bound=2)
    host_s = time.perf_counter() - t0
    lidx = np.repeat(np.arange(l_max + 1), np.arange(1, l_max + 2)).astype(np.int32)
    L = len(lidx)
    rng = np.random.default_rng(NOISE["seed"])
    rhs = torch.as_tensor(rng.standard_normal((L, n)) + 1j * rng.standard_normal((L, n))).cuda()
    d_l = torch.as_tensor(lidx).cuda()
    rows = torch.argsort(d_l.to(torch.int64), stable=True).to(torch.int32)
    out = torch.empty_like(rhs)
    ev = lambda: torch.cuda.Event(enable_timing=True)
    ops.factor()                                  # warm-up (module load, first launch)
    ops.solve("W", rhs, d_l, out, rows)
    torch.cuda.synchronize()
    e0, e1 = ev(), ev()
    e0.record()
    ops.factor()
    e1.record()
    solve_ms = {}
    for name in ("S", "Z", "W"):
        a, b = ev(), ev()
        a.record()
        ops.solve(name, rhs, d_l, out, rows)
        b.record()
        solve_ms[name] = (a, b)
    torch.cuda.synchronize()
    fac_ms = e0.elapsed_time(e1)
    sm = {k: a.elapsed_time(b) for k, (a, b) in solve_ms.items()}
    bytes_solve = 2.0 * L * n * 16.0                # right-hand sides in, solutions out
    return dict(workload="FD4 n_r=512 (stretched), l_max=255, polytropic N_rho=3 n=2; "
                         "S, Z kl=ku=2, W kl=ku=3; 32896 complex RHS per system",
                host_build_s=host_s, factor_ms_all_systems=fac_ms, solve_ms=sm,
                solve_GBps={k: bytes_solve / (v * 1e-3) / 1e9 for k, v in sm.items()},
                info_ok=all(int(ops.dev[k]["info"].abs().sum()) == 0 for k in ("S", "Z", "W")))


def sht_case():
    """Config 2: SHT round trip at l_max=383 (n_theta=576, n_phi=1152) on 3 x 97 = 291
    fields with seeded Gaussian coefficients scaled by (l+1)^-2 (m=0 real), in FP64 and
    in mixed precision: ms per direction (synthesis + inverse FFT; forward FFT +
    analysis) and the max / RMS round-trip error relative to the largest coefficient
    (analysis(synthesis(x)) = x exactly for band-limited x), plus the mixed grid
    against the FP64 grid."""
    import numpy as np
    import torch
    from magic_amd.device import Resolution
    from magic_amd.params import Params
    from magic_amd.sht import SHT
    p = Params({"l_max": 383, "n_r_max": 97})
    res = Resolution(p)
    B = 3 * 97
    rng = np.random.default_rng(NOISE["seed"])
    l, m = res.maps["l"], res.maps["m"]
    S = (rng.standard_normal((p.lm_max, B)) + 1j * rng.standard_normal((p.lm_max, B))) \
        * ((l + 1.0) ** -2.0)[:, None]
    S[m == 0] = S[m == 0].real
    dS = torch.from_numpy(S).cuda()
    smax = np.abs(S).max()
    ev = lambda: torch.cuda.Event(enable_timing=True)
    out, grids = {}, {}
    flops = 4.0 * p.lm_max * (p.n_theta // 2) * B
    for prec in ("fp64",) + MIXED:
        sht = SHT(res, precision=prec)
        g = sht.scal_to_spat(dS)
        back = sht.scal_to_SH(g)                       # warm-up (tables, module load)
        torch.cuda.synchronize()
        t = {"syn": [], "ana": []}
        for _ in range(3):
            a, b, c = ev(), ev(), ev()
            a.record()
            g = sht.scal_to_spat(dS)
            b.record()
            back = sht.scal_to_SH(g)
            c.record()
            torch.cuda.synchronize()
            t["syn"].append(a.elapsed_time(b))
            t["ana"].append(b.elapsed_time(c))
        err = (back - dS).abs().cpu().numpy()
        out[prec] = dict(synthesis_ms=float(np.min(t["syn"])), analysis_ms=float(np.min(t["ana"])),
                         roundtrip_max_err=float(err.max() / smax),
                         roundtrip_rms_err=float(np.sqrt((err ** 2).mean()) / smax),
                         legendre_tflops_synthesis_upper=flops / (np.min(t["syn"]) * 1e-3) / 1e12)
        grids[prec] = g
    gmax = grids["fp64"].abs().max()
    for prec in MIXED:
        d = (grids[prec] - grids["fp64"]).abs()
        out[f"{prec}_grid_vs_fp64"] = dict(max_err=float(d.max() / gmax),
                                           rms_err=float(d.pow(2).mean().sqrt() / gmax))
    out["workload"] = "l_max=383, 576 x 1152 grid, 291 fields, coefficients N(0,1) (l+1)^-2"
    return out


def ensemble_case(n_seq):
    """Config 5: 64 Boussinesq runs (l_max in {32, 64, 85, 127}, n_r in {33, 49, 65},
    E log-uniform [1e-4, 1e-3], Ra_mod uniform [60, 200], 500 steps each) arriving as a
    seeded Poisson process at 2/s, served with continuous admission (16 slots, HBM
    budget); baseline: the first n_seq of the same requests run back to back, one at
    a time, in plain FP64 (arrivals ignored)."""
    import numpy as np
    from magic_amd.serve import EnsembleServer, config5_requests, serve_sequential
    cfgs = [r.cfg for r in config5_requests()]
    srv = EnsembleServer(max_slots=16)
    tw = time.perf_counter()
    srv.warm(cfgs)                 # device tables of the served resolutions, before serving
    warm_s = time.perf_counter() - tw
    _, mix = srv.serve(config5_requests(), realtime=True)
    # the same 64 runs all waiting at t=0: throughput of the server when saturated
    sat_srv = EnsembleServer(max_slots=16)
    sat_srv.warm(cfgs)
    _, sat = sat_srv.serve(config5_requests(), realtime=False)
    _, seq = serve_sequential(config5_requests()[:n_seq], warm=True)
    return dict(workload="64 runs x 500 steps, Poisson arrivals 2/s, l_max {32,64,85,127}, "
                         "n_r {33,49,65}, FP64; each step one HIP graph with the dt decision "
                         "on the device, one step per run per round",
                table_warm_s=warm_s,
                server=dict(mix, max_active=max(n for _, n in srv.occupancy),
                            admit_ms_mean=1e3 * float(np.mean(srv.admit_s)),
                            admit_ms_max=1e3 * float(np.max(srv.admit_s)),
                            admit_stage_ms_max=1e3 * float(np.max(srv.stage_s))),
                saturated=dict(sat, max_active=max(n for _, n in sat_srv.occupancy)),
                sequential=dict(seq, sample=f"first {n_seq} of the 64 requests back to back"))


def measured_peaks():
    """Peaks measured on this device, beside the published ones: HBM from a
    2 GiB device-to-device copy (bytes read + written), FP64 and FP32 matrix
    rates from 8192^3 library GEMMs (torch.matmul, TF32 off); best of the
    repetitions, timed with HIP events on the current stream."""
    import torch
    dev = torch.device("cuda")

    def best(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t = float("inf")
        for _ in range(reps):
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            t = min(t, e0.elapsed_time(e1) * 1e-3)
        return t

    out = {}
    n = (2 << 30) // 8
    a = torch.ones(n, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    out["hbm_copy_GBs"] = 2 * n * 8 / best(lambda: b.copy_(a), 10) / 1e9
    del a, b
    tf32 = torch.backends.cuda.matmul.allow_tf32
    torch.backends.cuda.matmul.allow_tf32 = False
    m = 8192
    for name, dt in (("fp64", torch.float64), ("fp32", torch.float32)):
        g = torch.Generator(device=dev).manual_seed(1)
        A = torch.randn(m, m, dtype=dt, device=dev, generator=g)
        B = torch.randn(m, m, dtype=dt, device=dev, generator=g)
        C = torch.empty(m, m, dtype=dt, device=dev)
        out[f"{name}_gemm_TFLOPs"] = 2 * m ** 3 / best(lambda: torch.matmul(A, B, out=C), 5) / 1e12
        del A, B, C
    torch.backends.cuda.matmul.allow_tf32 = tf32
    torch.cuda.empty_cache()
    out["method"] = ("2 GiB float64 copy_ (read + write bytes); 8192^3 torch.matmul in FP64 and "
                     "FP32 (TF32 off); best of 10 / 5 repetitions")
    return out


def roofline(res, precision):
    """Roofline of the dominant kernel section (largest total time)."""
    secs, work = res["secs"], section_work(res["p"], res["N"], 16.0 if precision == "fp64" else 8.0)
    dom = max(secs, key=lambda k: secs[k][0] * secs[k][1])
    cnt, ms = secs[dom]
    # every mixed variant runs the same FP32-buffer grid step
    tprec = "fp64" if precision == "fp64" else "mixed"
    tkey = (dom, tprec, res["grid_rb"]) if dom == "grid_step" else (dom, precision)
    kind, amount = work.get(dom, ("hbm", float("nan")))
    if kind == "mfma":
        peak = {"mixed": PEAK_FP32_TFLOPS, "mixed_bf16": PEAK_BF16_TFLOPS,
                "mixed_bf16x3": PEAK_BF16_TFLOPS}.get(precision, PEAK_FP64_TFLOPS) \
            if dom.startswith("leg_") else PEAK_FP64_TFLOPS
        ach = amount / (ms * 1e-3) / 1e12
        return dict(bound="mfma", achieved=ach, peak=peak, unit="TFLOP/s", frac=ach / peak,
                    traffic=PMC_TRAFFIC_BYTES.get(tkey), kernel=dom, launch_ms=ms,
                    algorithmic_flops=amount)
    ach = amount / (ms * 1e-3) / 1e9
    return dict(bound="hbm", achieved=ach, peak=PEAK_HBM_GBS, unit="GB/s", frac=ach / PEAK_HBM_GBS,
                traffic=PMC_TRAFFIC_BYTES.get(tkey), kernel=dom, launch_ms=ms,
                algorithmic_bytes=amount)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", choices=("both", "all") + MIXED + ("fp64",), default="both",
                    help="both: FP64 and the --mixed variant; all: FP64 and every mixed variant")
    ap.add_argument("--mixed", choices=MIXED, default=HEAD_MIXED,
                    help="the mixed-precision variant of the headline value")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fd", action="store_true", help="skip the config-3 banded-solve timing")
    ap.add_argument("--no-published", action="store_true",
                    help="skip the step times at the reference's published points")
    ap.add_argument("--fd-step-only", type=int, default=0, metavar="STEPS",
                    help="only time STEPS finite-difference steps at config 3's physics "
                         "(FD4, n_r=512, l_max=255, N_rho=3) and print their JSON")
    ap.add_argument("--cpu-levels", type=int, default=9)
    ap.add_argument("--no-ensemble", action="store_true", help="skip the config-5 ensemble")
    ap.add_argument("--no-sht", action="store_true", help="skip the config-2 SHT round trip")
    ap.add_argument("--no-peaks", action="store_true",
                    help="skip the measured HBM / GEMM peaks")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 256 MiB device copies that calibrate the PMC byte "
                         "counters, then exit")
    ap.add_argument("--ensemble-seq", type=int, default=16,
                    help="requests in the sequential ensemble baseline sample")
    args = ap.parse_args()
    if args.fd_step_only:
        print(json.dumps(fd_step_case(steps=args.fd_step_only, warmup=2)))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration()
        return
    import numpy as np
    precs = {"both": ("fp64", args.mixed), "all": ("fp64",) + MIXED}.get(args.precision,
                                                                          (args.precision,))
    res = {}
    for pr in precs:
        # the mixed run replays the FP64 run's step sizes, so the two trajectories
        # differ by arithmetic precision only (the Courant dt = 0.75 dt_min moves
        # with the last bits of the velocity maxima)
        sched = res["fp64"]["dt_log"] if (pr != "fp64" and "fp64" in res) else None
        res[pr] = run_case(pr, args.steps, args.warmup, sched)
        print(f"bench: {pr} {args.steps} steps in {res[pr]['elapsed']:.2f} s", file=sys.stderr,
              flush=True)
    head = args.mixed if args.mixed in res else next((m for m in MIXED if m in res), "fp64")
    h = res[head]
    sps = {pr: args.steps / r["elapsed"] for pr, r in res.items()}
    breakdown = {pr: {k: round(v[1], 4) for k, v in sorted(r["secs"].items(), key=lambda kv: -kv[1][1])}
                 for pr, r in res.items()}
    cfg = {"workload": "Boussinesq CNAB2 dynamo step, l_max=383, n_r_max=97, n_theta=576, "
                       "n_phi=1152, E=1e-5, Ra=2e8, Pr=1, Pm=1",
           "l_max": 383, "n_r_max": 97, "precision": head,
           "precision_split": (LEG_SPLIT[head] + "; Fourier buffers, azimuthal FFTs, nonlinear "
                               "products and Courant maxima FP32; implicit solves, radial "
                               "derivatives, IMEX terms and energy/spectrum diagnostics FP64")
                              if head != "fp64" else "all FP64",
           "steps_per_s": {pr: v for pr, v in sps.items()},
           "ms_per_step": {pr: 1e3 * r["elapsed"] / args.steps for pr, r in res.items()},
           "dt": h["dt"], "rebuilds": {pr: r["rebuilds"] for pr, r in res.items()},
           "grid_levels_per_workgroup": {pr: r["grid_rb"] or "plan default" for pr, r in res.items()},
           "synthesis_waves_and_column_tiles": {pr: {"vector": list(r["syn_shape"][0]),
                                                     "scalar": list(r["syn_shape"][1])}
                                                for pr, r in res.items()},
           "row_operator_tile": {pr: "32x128" if r["rowop_wide"] else "32x64" for pr, r in res.items()},
           "analysis_waves_per_workgroup": {pr: r["ana_waves"] or "default" for pr, r in res.items()},
           "energies": {pr: r["energies"] for pr, r in res.items()},
           "section_ms": breakdown}
    if "fp64" in res and head != "fp64":
        cfg["mixed_over_fp64"] = sps[head] / sps["fp64"]
        e64 = np.array(res["fp64"]["energies"])
        per = {}
        for mv in [m for m in MIXED if m in res]:
            emx = np.array(res[mv]["energies"])
            sp = {}
            for k in ("kin_l", "mag_l", "kin_m", "mag_m"):
                a, b = res["fp64"]["spectra"][k], res[mv]["spectra"][k]
                sp[k] = float(np.max(np.abs(b - a)) / np.max(np.abs(a)))
            per[mv] = dict(energy_rel_diff=float(np.max(np.abs(emx - e64) / np.abs(e64))),
                           spectrum_rel_diff=sp, over_fp64=sps[mv] / sps["fp64"])
        cfg["energy_rel_diff_mixed_vs_fp64"] = per[head]["energy_rel_diff"]
        cfg["spectrum_rel_diff_mixed_vs_fp64"] = per[head]["spectrum_rel_diff"]
        cfg["energy_rel_diff_note"] = (f"after {args.warmup + args.steps} steps, mixed run on the "
                                       "FP64 run's dt sequence; max over e_kin_pol/tor, e_mag_pol/tor")
        if len(per) > 1:
            cfg["mixed_variants"] = per
    if not args.no_sht:
        cfg["sht_roundtrip"] = sht_case()
        print("bench: SHT round trip done", file=sys.stderr, flush=True)
    if not args.no_fd:
        cfg["fd_banded"] = fd_banded_case()
        print("bench: FD banded solves done", file=sys.stderr, flush=True)
        cfg["fd_step"] = fd_step_case()
        print(f"bench: FD step {cfg['fd_step']['ms_per_step']:.2f} ms", file=sys.stderr,
              flush=True)
    if not args.no_published:
        cfg["published_points"] = published_case()
    if not args.no_ensemble:
        cfg["ensemble"] = ensemble_case(args.ensemble_seq)
        print(f"bench: ensemble wall {cfg['ensemble']['server']['wall_s']:.1f} s", file=sys.stderr,
              flush=True)
    rl = roofline(h, head)
    if not args.no_peaks:
        pk = measured_peaks()
        cfg["measured_peaks"] = pk
        meas = {"GB/s": pk["hbm_copy_GBs"],
                "TFLOP/s": pk["fp32_gemm_TFLOPs" if head == "mixed" and rl["kernel"].startswith("leg_")
                              else "fp64_gemm_TFLOPs"]}[rl["unit"]]
        if rl["unit"] == "TFLOP/s" and head in ("mixed_bf16", "mixed_bf16x3") and \
                rl["kernel"].startswith("leg_"):
            meas = float("nan")   # no BF16 GEMM is measured
        rl["frac_of_measured_peak"] = rl["achieved"] / meas
        print("bench: measured peaks done", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_levels)
        print(f"bench: CPU baseline step {cpu['seconds']:.1f} s (n_r_max={args.cpu_levels}), "
              f"extrapolated {cpu['extrapolated_seconds_per_step_at_97']:.0f} s at 97",
              file=sys.stderr, flush=True)
    line = {
        "metric": "steps/s at l_max=383 mixed vs FP64; ensemble p50/p99 step latency",
        "value": sps[head],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * h["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[head],
        "data": "synthetic (seeded noise + dipole seed, generated on device)",
        "config": cfg,
        "roofline": rl,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
