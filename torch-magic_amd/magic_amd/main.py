"""The run driver and restart entry points under the reference's module name
(``main.py``), acting on the configured default run (``config.configure``).

``run(cfg)`` is ``main.run`` (reference ``main.py:280-993``): configuration ->
fresh start (``initialize_fields``, ``setup_initial_state``,
``initialize_dt``), restart from a ``.pt`` file, or a fresh start from a
Fortran MagIC checkpoint (``fortran_restart``) -> the step loop, with the
reference's log cadence (every ``log_every`` steps and the last) and
checkpoint cadence (``checkpoint_every``), and its files in ``output_dir``:
``log.csv`` (same columns), ``e_kin.TAG``, ``e_mag_oc.TAG``, ``timestep.TAG``
(a line at the start and at every dt change), ``radius.TAG``, ``signal.TAG``
and ``checkpoint_NNNNNN.pt``; at the end ``G_1.TAG`` (``n_graphs`` >= 1) and
the Fortran-format ``checkpoint_end.TAG``.  Returns (e_kin_pol, e_kin_tor,
e_mag_pol, e_mag_tor) like the reference's ``_get_energies``.

The steps between two log or checkpoint points run with the dt decision on
the device, one HIP graph per step (``DynamoRun.run_steps``): the same steps,
bit for bit, as calling ``one_step`` in a loop.  Not written: TensorBoard
scalars (the reference's ``SummaryWriter``), the par / dipole / heat / power /
helicity / hemi files, radial profile files and the log.TAG narrative.
Movies (``movies``, ``n_movie_step``) are written as the reference writes them:
the frame of step k holds the state before step k.
"""
from __future__ import annotations

import csv
import os
import time as _time

from . import config
from .checkpoint import load_checkpoint as _load, save_checkpoint as _save


def save_checkpoint(path, step, sim_time, cfg=None):
    """Write the default run's restart file (reference ``main.save_checkpoint``;
    the file's ``cfg`` entry is the run's own configuration)."""
    _save(config.default_run(), path, step, sim_time)


def load_checkpoint(path):
    """Restore the default run from a restart file; returns (step, sim_time)
    (reference ``main.load_checkpoint``)."""
    return _load(config.default_run(), path)


_CSV_COLUMNS = ["step", "sim_time", "e_kin_pol", "e_kin_tor", "e_mag_pol", "e_mag_tor",
                "wall_elapsed", "ms_per_step"]


def _energies(run):
    e = run.energies()
    return e if run.p.l_mag else (e[0], e[1], 0.0, 0.0)


def run(cfg=None, on_log=None):
    """Run the configured dynamo (reference ``main.run``, ``main.py:280-993``).

    cfg: the YAML dict (run keys ``n_steps``, ``dt``, ``log_every``,
    ``checkpoint_every``, ``output_dir``, ``restart``, ``n_graphs``,
    ``movies``, ``n_movie_step``, ``tag``); when nothing was configured yet,
    the same dict configures the run.  on_log: called after every log step."""
    from . import output
    cfg = dict(cfg or {})
    if "dt" in cfg and "dtmax" not in cfg:
        cfg["dtmax"] = cfg["dt"]
    if not config._config and cfg:
        config.configure(cfg)
    r = config.default_run()
    p = r.p
    n_steps = int(cfg.get("n_steps", 1000))
    dt = float(cfg.get("dt", p.dtmax))
    log_every = int(cfg.get("log_every", 100))
    checkpoint_every = int(cfg.get("checkpoint_every", 0))
    output_dir = cfg.get("output_dir", "./output")
    restart = cfg.get("restart", None)
    fortran_restart = cfg.get("fortran_restart", None)
    n_graphs = int(cfg.get("n_graphs", 1))
    movies = cfg.get("movies", []) or []
    n_movie_step = int(cfg.get("n_movie_step", 0))
    if not (n_movie_step > 0 and movies):
        n_movie_step, movies = 0, []
    if movies and int(getattr(p, "radial_chunk_size", 0) or 0) > 0:
        print("  Warning: movies disabled in chunked radial mode")   # as main.py:465-468
        n_movie_step, movies = 0, []
    tag = cfg.get("tag", "torch")
    os.makedirs(output_dir, exist_ok=True)

    if fortran_restart:
        from .checkpoint_io import restart_from_fortran
        sim_time, dt = restart_from_fortran(r, fortran_restart)
        start_step = 0
    elif restart:
        start_step, sim_time = r.load_checkpoint(restart)
        dt = r.dt_hist[0]
    else:
        start_step, sim_time = 0, 0.0
        r.initialize_fields()
        r.setup_initial_state()
        r.initialize_dt(dt)   # the loop keeps passing the dt given, as the reference does
    appending = bool(restart) and os.path.exists(os.path.join(output_dir, "log.csv"))
    mode = "a" if appending else "w"
    path = lambda name: os.path.join(output_dir, f"{name}.{tag}")
    f_csv = open(os.path.join(output_dir, "log.csv"), mode, newline="")
    f_ekin = open(path("e_kin"), mode)
    f_emag = open(path("e_mag_oc"), mode) if p.l_mag else None
    f_ts = open(path("timestep"), mode)
    files = [f for f in (f_csv, f_ekin, f_emag, f_ts) if f is not None]
    mov = None
    if movies:
        from .movie_output import MovieWriter
        mov = MovieWriter(r, movies, output_dir, tag)   # new files, also on restart
    try:
        writer = csv.writer(f_csv)
        if not appending:
            writer.writerow(_CSV_COLUMNS)
            output.write_radius_file(path("radius"), r.res.rad["r"])
            output.write_signal_file(path("signal"))
            output.write_timestep_line(f_ts, 0.0, dt)
            f_ts.flush()

        def write_row(step, t, wall, steps_done):
            ek, em = r.energy_components()
            e = (ek[0], ek[1], em[0], em[1]) if p.l_mag else (ek[0], ek[1], 0.0, 0.0)
            ms = wall / steps_done * 1000.0 if steps_done > 0 else 0.0
            writer.writerow([step, f"{t:.10e}", *(f"{v:.15e}" for v in e), f"{wall:.3f}",
                             f"{ms:.3f}"])
            f_csv.flush()
            output.write_e_kin_line(f_ekin, t, output.EKin(*ek))
            f_ekin.flush()
            if f_emag is not None:
                output.write_e_mag_oc_line(f_emag, t, output.EMagOC(*em))
                f_emag.flush()
            print(f"Step {step:6d}: e_kin={e[0] + e[1]:.6e} (pol={e[0]:.6e} tor={e[1]:.6e}) "
                  f"e_mag={e[2] + e[3]:.6e} (pol={e[2]:.6e} tor={e[3]:.6e}) "
                  f"dt={dt:.6e} t={t:.6e} wall={wall:.1f}s")

        if not restart:
            write_row(0, sim_time, 0.0, 0)
        r.use_graphs = True
        t0 = _time.time()
        n = start_step
        while n < n_steps:
            # next stop: a log step, a checkpoint step or the last step
            stop = n_steps
            if log_every > 0:
                stop = min(stop, (n // log_every + 1) * log_every)
            if checkpoint_every > 0:
                stop = min(stop, (n // checkpoint_every + 1) * checkpoint_every)
            movie_step = False
            if mov is not None:
                # a movie frame at step k holds the grids of step k's radial loop,
                # i.e. the state before step k (main.py:852-868): take it, run the step
                if (n + 1) % n_movie_step == 0:
                    mov.capture()
                    stop, movie_step = n + 1, True
                else:
                    stop = min(stop, (n // n_movie_step + 1) * n_movie_step - 1)
            k0 = len(r.dt_log)
            r.run_steps(stop - n, dt)
            for dt_actual in r.dt_log[k0:]:
                if dt_actual != dt:
                    output.write_timestep_line(f_ts, sim_time + dt_actual, dt_actual)
                    f_ts.flush()
                dt = dt_actual
                sim_time += dt
            n = stop
            if movie_step:
                mov.write(sim_time)
            elapsed = _time.time() - t0
            if (log_every > 0 and n % log_every == 0) or n == n_steps:
                write_row(n, sim_time, elapsed, n - start_step)
                if on_log:
                    on_log()
            if checkpoint_every and (n % checkpoint_every == 0 or n == n_steps):
                cp = os.path.join(output_dir, f"checkpoint_{n:06d}.pt")
                r.save_checkpoint(cp, n, sim_time)
                print(f"  Saved checkpoint: {cp}")
        if n_steps > start_step:
            if n_graphs >= 1:
                from .graph_output import write_graph_file
                write_graph_file(path("G_1"), sim_time, r)
            from .checkpoint_io import write_checkpoint_fortran
            write_checkpoint_fortran(path("checkpoint_end"), r, sim_time, n_steps + 1)
            done = n_steps - start_step
            elapsed = _time.time() - t0
            print(f"\nDone: {done} steps in {elapsed:.1f}s ({elapsed / done * 1000:.1f}ms/step)")
    finally:
        for f in files:
            f.close()
        if mov is not None:
            mov.close()
    return _energies(r)
