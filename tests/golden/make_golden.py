"""Write golden fixtures by running the reference on the CPU.

What closed forms and library routines cannot pin is generated here from the
reference package itself (``/root/reference/src``, pure Python, FP64,
``MAGIC_DEVICE=cpu``), one child process per case, and stored as small
``.npz`` vectors beside this script.  It is run by hand in the development
container; tests only read the ``.npz`` files.

Trajectory cases (``step_time.one_step``, ``step_time.py:1574-1637``: radial
loop, CFL, IMEX weights, rebuild on a dt change, fused lm loop, insulating B
update).  Each: ``initialize_fields`` -> optional seeded noise on s ->
``setup_initial_state`` -> ``initialize_dt`` -> ``one_step`` x n, with the dt
that each step returns passed to the next (as ``main.py`` does):

* ``bench16``: dynamo benchmark l_max=16, n_r_max=33, no noise, 20 steps;
  energies every step, fields at steps 1 and 20, explicit terms after the
  first radial loop.
* ``c1``: l_max=32, n_r_max=33, E=1e-3, Ra=1e5, Pr=1, Pm=5, dt=1e-4,
  noise 1e-5 (seed 20261202), 100 steps; energies every step, fields at
  step 100, and at step 100 the full e_kin / e_mag_oc tuples
  (``output.py:111-193``) and the length scales of ``get_dlm``
  (``output.py:750-861``) for the flow ('V') and the field ('B').
* ``dtchange``: l_max=16, n_r_max=33, dtmax=1e-3, noise 1.0 (seed 7), 30
  steps: the Courant limit changes dt several times, which exercises
  ``courant.py:80-155`` (``dt_courant``), the variable-step CNAB2 weights
  (``time_scheme.py:56-73``) and the rebuild at ``step_time.py:1628-1629``;
  energies and dt every step, fields at steps 10 and 30.

* ``bpr16`` and ``bprdt``: the BPR353 four-stage DIRK step
  (``time_scheme.py:157-326``, ``step_time.py:1640-1680``) on the dynamo
  benchmark (l_max=16, n_r_max=33; noise 1e-3 seed 11, dt 1e-4, 20 steps) and on
  the dt-change case above (dtmax 1e-3, noise 1.0 seed 7, 30 steps), whose
  Courant limit (BPR353's courfac 0.8, alffac 0.35) changes dt at stage 1.

* ``minc2`` and ``bprminc4``: azimuthal symmetry (``params.py:60-76``: orders
  m = 0, minc, 2 minc, ..., one 2 pi / minc sector of n_phi_tot / minc
  longitudes) with CNAB2 (l_max=16, minc=2) and with BPR353 (l_max=32, minc=4),
  n_r_max=33, 20 steps.

* ``condic`` and ``bprcondic``: a conducting inner core (sigma_ratio = 1,
  kbotb = 3, not rotating; the coupled outer/inner B and J solves of
  ``update_b.py:301-513,634-801`` on the even-Chebyshev inner grid of
  ``radial_functions.py:209-306``), l_max=16, n_r_max=33 with the benchmark's
  truncations n_cheb_max=31, n_cheb_ic_max=15, CNAB2 and BPR353, 20 steps;
  the inner-core fields b_ic, aj_ic at the saved steps.
* ``rotic`` and ``bprrotic4``: the same core rotating (nRotIC = 1, no-slip ICB:
  the z10 system of ``update_z.py:244-298,355-501``, the Lorentz torque and
  the rigid-rotation induction and advection of ``step_time.py:1126-1200``,
  ``update_b.py:875-901``), CNAB2 at l_max=16 and BPR353 at l_max=32 with
  minc=4 (the boussBenchSat combination), dt 2e-4, noise 1e-2, 30 steps;
  omega_ic every step.

* ``fd2mhd``, ``fd4mhd``, ``fdanel``, ``fdanelmhd``, ``fd4anel``, ``fdbpr``: the
  finite-difference radial scheme (stretched grid, fd_stretch 0.3, fd_ratio
  0.1) with the double-curl poloidal equation (``update_w_doublecurl.py``,
  ``get_td.py:91-147``): the physics of the reference's FD step tests
  (``tests/test_fd_step.py``, ``test_fd4_step.py``, ``test_fd_anel.py``,
  ``test_fd_anel_mhd.py``) with seeded noise, FD2 and FD4 (n_r even in
  ``fd4mhd``), config 3's anelastic physics (FD4, strat 3, polind 2) at
  l_max=16, n_r=48 in ``fd4anel``, a conducting and a rotating inner core
  (``fdcondic``, ``fdrotic``: the reference's bordered band solves,
  ``algebra.py:513-617``) and BPR353 in ``fdbpr``; 20 steps (10 for BPR353).

Finite-difference operator cases (``fd4_ob2``, ``fd4_ob3``): the FD branches
of ``build_s_matrices`` (``update_s.py:66-188``), ``build_z_matrices``
(``update_z.py:89-240``) and ``build_w_matrices``
(``update_w_doublecurl.py:67-307``) at fd_order 4 with boundary order 2 or 3,
n_r_max=64, l_max=12, anelastic polytrope strat=3, polind=2, g2=1, called
with wimp_lin0 = 6e-5: the radial grid, the background profiles, and for each
system the LINPACK-factored bands, pivots and preconditioning factors.

Output-file case (``files16``): l_max=16, n_r_max=17 (the config-1 physics,
noise 1e-3 seed 3, dt 1e-4), 10 steps, then the reference's own writers:
the Fortran MagIC v5 checkpoint (``checkpoint_io.write_checkpoint_fortran``,
``checkpoint_io.py:336-490``) and the graph file (``graph_output.
write_graph_file``, ``graph_output.py:53-242``), stored as written
(``files16_checkpoint.bin``, ``files16_G_1.bin``), with the state they were
written from (``files16.npz``: fields, explicit history slot, dt, time) and the
energies of 10 further steps continued in memory.  A second process restarts
from the checkpoint file the way ``main.run`` does for ``fortran_restart``
(``main.py:318-374``) and steps 10 times: ``energies_restart`` (the reference's
reader keeps the file's big-endian byte order, which its own
``load_fortran_checkpoint`` cannot hand to ``torch.from_numpy``; the script
gives it native-order copies of the same values).

Movie case (``movies16``): the reference's movie writer
(``movie_output.py:120-287``: header, two frames) on the grid fields of the
``files16`` state, computed as its radial loop does (``torpol_to_spat`` of
(l(l+1) w, dw, z) and (l(l+1) b, db, aj), ``scal_to_spat`` of s,
``step_time.py:955-1059``), for CMB and equatorial movies of Br, Bt, Vr, Vp and T
(``files16_<Field>_<CMB|EQU>_mov.bin``).

Usage: ``python tests/golden/make_golden.py [case ...]``
"""

from __future__ import annotations

import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_SRC = "/root/reference/src"

CASES = {
    "bench16": dict(cfg=dict(l_max=16, n_r_max=33), dt=1e-4, steps=20, noise=None,
                    field_steps=[1, 20], expl_step1=True),
    "c1": dict(cfg=dict(l_max=32, n_r_max=33, ek=1e-3, ra=1e5, pr=1.0, prmag=5.0,
                        radratio=0.35, dtmax=1e-4), dt=1e-4, steps=100,
               noise=dict(seed=20261202, amp=1e-5), field_steps=[100], expl_step1=False,
               diag_steps=[100]),
    "dtchange": dict(cfg=dict(l_max=16, n_r_max=33, ek=1e-3, ra=1e5, pr=1.0, prmag=5.0,
                              radratio=0.35, dtmax=1e-3), dt=1e-3, steps=30,
                     noise=dict(seed=7, amp=1.0), field_steps=[10, 30], expl_step1=False,
                     diag_steps=[30]),
    "bpr16": dict(cfg=dict(l_max=16, n_r_max=33, time_scheme="BPR353"), dt=1e-4, steps=20,
                  noise=dict(seed=11, amp=1e-3), field_steps=[1, 20], expl_step1=False),
    "bprdt": dict(cfg=dict(l_max=16, n_r_max=33, ek=1e-3, ra=1e5, pr=1.0, prmag=5.0,
                           radratio=0.35, dtmax=1e-3, time_scheme="BPR353"), dt=1e-3, steps=30,
                  noise=dict(seed=7, amp=1.0), field_steps=[10, 30], expl_step1=False),
    "minc2": dict(cfg=dict(l_max=16, n_r_max=33, minc=2), dt=1e-4, steps=20,
                  noise=dict(seed=13, amp=1e-3), field_steps=[1, 20], expl_step1=False,
                  diag_steps=[20]),
    "bprminc4": dict(cfg=dict(l_max=32, n_r_max=33, minc=4, time_scheme="BPR353"), dt=1e-4,
                     steps=20, noise=dict(seed=17, amp=1e-3), field_steps=[1, 20],
                     expl_step1=False, diag_steps=[20]),
    "condic": dict(cfg=dict(l_max=16, n_r_max=33, n_cheb_max=31, n_cheb_ic_max=15,
                            sigma_ratio=1.0, kbotb=3), dt=1e-4, steps=20,
                   noise=dict(seed=19, amp=1e-3), field_steps=[1, 20], expl_step1=False,
                   diag_steps=[20]),
    "bprcondic": dict(cfg=dict(l_max=16, n_r_max=33, n_cheb_max=31, n_cheb_ic_max=15,
                               sigma_ratio=1.0, kbotb=3, time_scheme="BPR353"), dt=1e-4,
                      steps=20, noise=dict(seed=23, amp=1e-3), field_steps=[1, 20],
                      expl_step1=False),
    "rotic": dict(cfg=dict(l_max=16, n_r_max=33, n_cheb_max=31, n_cheb_ic_max=15,
                           sigma_ratio=1.0, kbotb=3, nRotIC=1), dt=2e-4, steps=30,
                  noise=dict(seed=29, amp=1e-2), field_steps=[1, 30], expl_step1=False),
    "bprrotic4": dict(cfg=dict(l_max=32, n_r_max=33, n_cheb_max=31, n_cheb_ic_max=15, minc=4,
                               sigma_ratio=1.0, kbotb=3, nRotIC=1, time_scheme="BPR353"),
                      dt=2e-4, steps=30, noise=dict(seed=31, amp=1e-2), field_steps=[1, 30],
                      expl_step1=False),
    # anelastic (polytropic) step: the hydrodynamic benchmark's physics at small
    # size (stress-free walls, angular-momentum corrections, (19, 19) seed), a
    # magnetic anelastic run (gravity ~ r, no-slip), and BPR353 with minc = 2
    "anel21": dict(cfg=dict(l_max=21, n_r_max=33, n_cheb_max=31, mode=1, strat=5.0,
                            polind=2.0, g0=0.0, g1=0.0, g2=1.0, ra=1.48638035e5, ek=1e-3,
                            pr=1.0, ktopv=1, kbotv=1, l_correct_AMz=True, l_correct_AMe=True,
                            init_s1=1919, amp_s1=0.01, alpha=0.6, dtmax=1e-4),
                   dt=1e-4, steps=30, noise=dict(seed=43, amp=1e-3), field_steps=[1, 30],
                   expl_step1=True, diag_steps=[30]),
    "anelmag16": dict(cfg=dict(l_max=16, n_r_max=33, strat=1.0, polind=2.0, g1=1.0, ra=1e5,
                               ek=1e-3, pr=1.0, prmag=5.0, dtmax=1e-4),
                      dt=1e-4, steps=20, noise=dict(seed=47, amp=1e-3), field_steps=[1, 20],
                      expl_step1=True, diag_steps=[20]),
    "bpranel2": dict(cfg=dict(l_max=16, n_r_max=33, n_cheb_max=31, minc=2, mode=1, strat=3.0,
                              polind=2.0, g2=1.0, ra=1e5, ek=1e-3, pr=1.0, ktopv=1, kbotv=1,
                              l_correct_AMz=True, dtmax=1e-4, time_scheme="BPR353"),
                     dt=1e-4, steps=20, noise=dict(seed=53, amp=1e-3), field_steps=[1, 20],
                     expl_step1=False),
    # rotating inner core started at omega_ic1 = 2 (initV, init_fields.py:528-532)
    "rotom8": dict(cfg=dict(l_max=8, n_r_max=17, n_cheb_max=15, n_cheb_ic_max=7,
                            sigma_ratio=1.0, kbotb=3, nRotIC=1, omega_ic1=2.0, dtmax=2e-4),
                   dt=2e-4, steps=15, noise=dict(seed=61, amp=1e-2), field_steps=[1, 15],
                   expl_step1=False),
    # Boussinesq, stress-free walls with both angular-momentum corrections
    "amc16": dict(cfg=dict(l_max=16, n_r_max=33, ktopv=1, kbotv=1, l_correct_AMz=True,
                           l_correct_AMe=True, dtmax=1e-4),
                  dt=1e-4, steps=20, noise=dict(seed=59, amp=1e-2), field_steps=[1, 20],
                  expl_step1=False),
    # finite-difference radial scheme (double-curl poloidal equation): the physics
    # of the reference's FD step tests (test_fd_step.py, test_fd4_step.py,
    # test_fd_anel.py, test_fd_anel_mhd.py) with seeded noise
    "fd2mhd": dict(cfg=dict(l_max=16, n_r_max=33, radial_scheme="FD", fd_order=2,
                            fd_order_bound=2, fd_stretch=0.3, fd_ratio=0.1, ra=1e5, ek=1e-3,
                            pr=1.0, prmag=5.0, dtmax=1e-4, alpha=0.6, init_s1=404, amp_s1=0.1,
                            ktopv=2, kbotv=2),
                   dt=1e-4, steps=20, noise=dict(seed=67, amp=1e-3), field_steps=[1, 20],
                   expl_step1=True, diag_steps=[20]),
    "fd4mhd": dict(cfg=dict(l_max=16, n_r_max=40, radial_scheme="FD", fd_order=4,
                            fd_order_bound=2, fd_stretch=0.3, fd_ratio=0.1, ra=1e5, ek=1e-3,
                            pr=1.0, prmag=5.0, dtmax=1e-4, alpha=0.6, init_s1=404, amp_s1=0.1,
                            ktopv=2, kbotv=2),
                   dt=1e-4, steps=20, noise=dict(seed=71, amp=1e-3), field_steps=[1, 20],
                   expl_step1=False, diag_steps=[20]),
    "fdanel": dict(cfg=dict(l_max=16, n_r_max=33, radial_scheme="FD", fd_order=2,
                            fd_order_bound=2, fd_stretch=0.3, fd_ratio=0.1, strat=0.1,
                            polind=2.0, g0=0.0, g1=0.0, g2=1.0, mode=1, ktopv=1, kbotv=1,
                            ra=1.48638035e5, ek=1e-3, pr=1.0, dtmax=1e-4, alpha=0.6,
                            init_s1=1010, amp_s1=0.01),
                   dt=1e-4, steps=20, noise=dict(seed=73, amp=1e-3), field_steps=[1, 20],
                   expl_step1=True, diag_steps=[20]),
    "fdanelmhd": dict(cfg=dict(l_max=16, n_r_max=33, radial_scheme="FD", fd_order=2,
                               fd_order_bound=2, fd_stretch=0.3, fd_ratio=0.1, strat=0.1,
                               polind=2.0, g0=0.0, g1=0.0, g2=1.0, mode=0, ktopv=2, kbotv=2,
                               ra=1e5, ek=1e-3, pr=1.0, prmag=5.0, dtmax=1e-4, alpha=0.6,
                               init_s1=404, amp_s1=0.1),
                      dt=1e-4, steps=20, noise=dict(seed=79, amp=1e-3), field_steps=[1, 20],
                      expl_step1=False),
    # config 3's physics (FD4, N_rho = 3, polytropic index 2, hydrodynamic) at a
    # small size
    "fd4anel": dict(cfg=dict(l_max=16, n_r_max=48, radial_scheme="FD", fd_order=4,
                             fd_order_bound=2, fd_stretch=0.3, fd_ratio=0.1, strat=3.0,
                             polind=2.0, g0=0.0, g1=0.0, g2=1.0, mode=1, ktopv=1, kbotv=1,
                             ra=1e5, ek=1e-3, pr=1.0, dtmax=1e-4, alpha=0.6, init_s1=1010,
                             amp_s1=0.01),
                    dt=1e-4, steps=20, noise=dict(seed=89, amp=1e-3), field_steps=[1, 20],
                    expl_step1=True, diag_steps=[20]),
    # conducting and rotating inner core on the finite-difference grid (the
    # physics of test_fd_condic_step.py and test_fd_rotic_step.py)
    "fdcondic": dict(cfg=dict(l_max=16, n_r_max=33, radial_scheme="FD", fd_order=2,
                              fd_order_bound=2, fd_stretch=0.3, fd_ratio=0.1, ra=1e5, ek=1e-3,
                              pr=1.0, prmag=5.0, dtmax=1e-4, alpha=0.6, init_s1=404,
                              amp_s1=0.1, ktopv=2, kbotv=2, kbotb=3, sigma_ratio=1.0,
                              n_cheb_ic_max=15),
                     dt=1e-4, steps=20, noise=dict(seed=97, amp=1e-3), field_steps=[1, 20],
                     expl_step1=False),
    "fdrotic": dict(cfg=dict(l_max=16, n_r_max=33, radial_scheme="FD", fd_order=2,
                             fd_order_bound=2, fd_stretch=0.3, fd_ratio=0.1, ra=1e5, ek=1e-3,
                             pr=1.0, prmag=5.0, dtmax=2e-4, alpha=0.6, init_s1=404,
                             amp_s1=0.1, ktopv=2, kbotv=2, kbotb=3, sigma_ratio=1.0,
                             n_cheb_ic_max=15, nRotIC=1),
                    dt=2e-4, steps=20, noise=dict(seed=101, amp=1e-2), field_steps=[1, 20],
                    expl_step1=False),
    "fdbpr": dict(cfg=dict(l_max=16, n_r_max=33, radial_scheme="FD", fd_order=2,
                           fd_order_bound=2, fd_stretch=0.3, fd_ratio=0.1, ra=1e5, ek=1e-3,
                           pr=1.0, prmag=5.0, dtmax=1e-4, init_s1=404, amp_s1=0.1,
                           time_scheme="BPR353"),
                  dt=1e-4, steps=10, noise=dict(seed=83, amp=1e-3), field_steps=[1, 10],
                  expl_step1=False),
}

FD_CASES = {
    "fd4_ob2": dict(cfg=dict(l_max=12, n_r_max=64, radial_scheme="FD", fd_order=4,
                             fd_order_bound=2, strat=3.0, polind=2.0, g0=0.0, g1=0.0, g2=1.0,
                             mode=1, ra=1e5, ek=1e-3, pr=1.0, radratio=0.35), wimp=6e-5),
    "fd4_ob3": dict(cfg=dict(l_max=12, n_r_max=64, radial_scheme="FD", fd_order=4,
                             fd_order_bound=3, strat=3.0, polind=2.0, g0=0.0, g1=0.0, g2=1.0,
                             mode=1, ra=1e5, ek=1e-3, pr=1.0, radratio=0.35), wimp=6e-5),
}

CHILD = r"""
import json, os, sys
import numpy as np
case = json.loads(sys.argv[1]); out = sys.argv[2]
sys.path.insert(0, sys.argv[3]); sys.path.insert(0, sys.argv[4])
from magic_torch.config import configure
configure(dict(case["cfg"], device="cpu", use_shtns=False))
import torch
torch.set_num_threads(8)
from magic_torch import fields, dt_fields, step_time, output
from magic_torch.init_fields import initialize_fields
initialize_fields()
if case["noise"]:
    from oracle.params import Params
    from oracle import horizontal, radial
    from oracle.dynamo import seeded_noise
    p = Params(case["cfg"])
    nz = seeded_noise(p, horizontal.lm_maps(p), radial.radial_functions(p),
                      case["noise"]["seed"], case["noise"]["amp"])
    fields.s_LMloc += torch.from_numpy(nz)
step_time.setup_initial_state()
step_time.initialize_dt(case["dt"])
names = ["s", "z", "w", "p", "b", "aj"]
res = {}
en = []
dts = []
dt = case["dt"]
for n in range(1, case["steps"] + 1):
    dt = step_time.one_step(n, dt)
    dts.append(dt)
    ek = output.get_e_kin_full(fields.w_LMloc, fields.dw_LMloc, fields.z_LMloc)
    em = output.get_e_mag_oc_full(fields.b_LMloc, fields.db_LMloc, fields.aj_LMloc)
    en.append([ek.e_p, ek.e_t, em.e_p, em.e_t])
    if int(case["cfg"].get("nRotIC", 0)) != 0:
        res.setdefault("omega_ic", []).append(float(fields.omega_ic))
    if n in case["field_steps"]:
        for k in names:
            res[f"{k}_step{n}"] = getattr(fields, k + "_LMloc").numpy().copy()
        if float(case["cfg"].get("sigma_ratio", 0.0)) > 0.0:
            for k in ("b_ic", "aj_ic"):
                res[f"{k}_step{n}"] = getattr(fields, k).numpy().copy()
        for k in ("dw", "db"):
            res[f"{k}_step{n}"] = getattr(fields, k + "_LMloc").numpy().copy()
    if n in case.get("diag_steps", []):
        res[f"ekin_full_step{n}"] = np.array(list(ek))
        res[f"emag_full_step{n}"] = np.array(list(em))
        for sw, (a, da, c) in (("V", (fields.w_LMloc, fields.dw_LMloc, fields.z_LMloc)),
                               ("B", (fields.b_LMloc, fields.db_LMloc, fields.aj_LMloc))):
            dl, dlR, dm, dlc, dlPolPeak, dlRc, dlPolPeakR = output.get_dlm(a, da, c, sw)
            res[f"dlm{sw}_scalars_step{n}"] = np.array([dl, dm, dlc, dlPolPeak])
            res[f"dlm{sw}_radial_step{n}"] = np.stack([dlR.numpy(), dlRc.numpy(),
                                                      dlPolPeakR.numpy()])
    if n == 1 and case["expl_step1"]:
        for k, ta in (("s", dt_fields.dsdt), ("z", dt_fields.dzdt), ("w", dt_fields.dwdt),
                      ("p", dt_fields.dpdt), ("b", dt_fields.dbdt), ("j", dt_fields.djdt)):
            # after the lm loop the step-1 explicit term sits in slot 1
            res[f"expl_{k}_step1"] = ta.expl[:, :, 1].numpy().copy()
res["energies"] = np.array(en)
res["dt"] = np.array(dts)
if "omega_ic" in res:
    res["omega_ic"] = np.array(res["omega_ic"])
np.savez(out, **res)
"""

FD_CHILD = r"""
import json, sys
import numpy as np
case = json.loads(sys.argv[1]); out = sys.argv[2]
sys.path.insert(0, sys.argv[3])
from magic_torch.config import configure
configure(dict(case["cfg"], device="cpu", use_shtns=False))
import torch
torch.set_num_threads(8)
from magic_torch import radial_functions as rf
from magic_torch import update_s, update_z, update_w_doublecurl as uw
res = {}
res["r"] = rf.r.numpy().copy()
for k in ("rho0", "temp0", "beta", "dbeta", "ddbeta", "dLtemp0", "orho1", "rgrav"):
    res["prof_" + k] = getattr(rf, k).numpy().copy()
w = case["wimp"]
update_s.build_s_matrices(w)
update_z.build_z_matrices(w)
uw.build_w_matrices(w)
res["S_abd"] = update_s._s_bands_by_l.numpy().copy()
res["S_piv"] = update_s._s_piv_by_l.numpy().copy()
res["S_fac"] = update_s._s_fac_by_l.numpy().copy()
res["S_k"] = np.array([update_s._s_kl, update_s._s_ku])
res["Z_abd"] = update_z._z_bands_by_l.numpy().copy()
res["Z_piv"] = update_z._z_piv_by_l.numpy().copy()
res["Z_fac"] = update_z._z_fac_by_l.numpy().copy()
res["Z_k"] = np.array([update_z._z_kl, update_z._z_ku])
res["W_abd"] = uw._w_bands_by_l.numpy().copy()
res["W_piv"] = uw._w_piv_by_l.numpy().copy()
res["W_fac_row"] = uw._w_fac_row_by_l.numpy().copy()
res["W_fac_col"] = uw._w_fac_col_by_l.numpy().copy()
res["W_k"] = np.array([uw._w_kl, uw._w_ku])
np.savez(out, **res)
"""


FILES_CASE = dict(cfg=dict(l_max=16, n_r_max=17, ek=1e-3, ra=1e5, pr=1.0, prmag=5.0,
                           radratio=0.35, dtmax=1e-4), dt=1e-4, steps=10,
                  noise=dict(seed=3, amp=1e-3), cont=10)

FILES_CHILD = r"""
import json, sys
import numpy as np
case = json.loads(sys.argv[1]); out = sys.argv[2]
sys.path.insert(0, sys.argv[3]); sys.path.insert(0, sys.argv[4])
from magic_torch.config import configure
configure(dict(case["cfg"], device="cpu", use_shtns=False))
import torch
torch.set_num_threads(8)
from magic_torch import fields, dt_fields, step_time, output
from magic_torch.init_fields import initialize_fields
from magic_torch.checkpoint_io import write_checkpoint_fortran
from magic_torch.graph_output import write_graph_file
from oracle.params import Params
from oracle import horizontal, radial
from oracle.dynamo import seeded_noise
initialize_fields()
p = Params(case["cfg"])
nz = seeded_noise(p, horizontal.lm_maps(p), radial.radial_functions(p),
                  case["noise"]["seed"], case["noise"]["amp"])
fields.s_LMloc += torch.from_numpy(nz)
step_time.setup_initial_state()
step_time.initialize_dt(case["dt"])
dt = case["dt"]; t = 0.0
for n in range(1, case["steps"] + 1):
    dt = step_time.one_step(n, dt)
    t += dt
write_checkpoint_fortran(out + "_checkpoint.bin", t, case["steps"] + 1)
write_graph_file(out + "_G_1.bin", t, fields.w_LMloc, fields.dw_LMloc, fields.z_LMloc,
                 fields.s_LMloc, fields.p_LMloc, fields.b_LMloc, fields.db_LMloc,
                 fields.aj_LMloc)
res = dict(time=np.array(t), dt=np.array([dt]))
for k in ("w", "dw", "z", "p", "s", "b", "db", "aj"):
    res[k] = getattr(fields, k + "_LMloc").numpy().copy()
for k, ta in (("w", dt_fields.dwdt), ("z", dt_fields.dzdt), ("p", dt_fields.dpdt),
              ("s", dt_fields.dsdt), ("b", dt_fields.dbdt), ("aj", dt_fields.djdt)):
    res["expl1_" + k] = ta.expl[:, :, 1].numpy().copy()
import magic_torch.time_scheme as ts
res["tscheme_dt"] = ts.tscheme.dt.numpy().copy()
en = []
for n in range(case["steps"] + 1, case["steps"] + case["cont"] + 1):
    dt = step_time.one_step(n, dt)
    ek = output.get_e_kin_full(fields.w_LMloc, fields.dw_LMloc, fields.z_LMloc)
    em = output.get_e_mag_oc_full(fields.b_LMloc, fields.db_LMloc, fields.aj_LMloc)
    en.append([ek.e_p, ek.e_t, em.e_p, em.e_t])
res["energies_cont"] = np.array(en)
np.savez(out + ".npz", **res)
"""

RESTART_CHILD = r"""
import json, sys
import numpy as np
case = json.loads(sys.argv[1]); out = sys.argv[2]
sys.path.insert(0, sys.argv[3])
from magic_torch.config import configure
configure(dict(case["cfg"], device="cpu", use_shtns=False))
import torch
torch.set_num_threads(8)
from magic_torch import fields, dt_fields, step_time, output
import magic_torch.main as _m
load_fortran_checkpoint = _m.load_fortran_checkpoint
from magic_torch.time_scheme import tscheme
from magic_torch.precision import CDTYPE, DEVICE
from magic_torch import checkpoint_io
# the reader keeps the file's (big-endian) byte order, which torch.from_numpy in
# load_fortran_checkpoint refuses: hand it native-order copies of the arrays
_read = checkpoint_io.read_checkpoint
def _native(path):
    ck = _read(path)
    nat = lambda a: a.astype(a.dtype.newbyteorder("=")) if isinstance(a, np.ndarray) else a
    for k in ("w", "z", "p", "s", "xi", "b", "aj", "b_ic", "aj_ic", "dt", "r"):
        setattr(ck, k, nat(getattr(ck, k)))
    ck.derivatives = {n: {part: [nat(a) for a in v] for part, v in d.items()}
                      for n, d in ck.derivatives.items()}
    return ck
checkpoint_io.read_checkpoint = _native
# main.run's fortran_restart branch (main.py:318-374), TensorBoard aside
ck = load_fortran_checkpoint(out + "_checkpoint.bin")
step_time.setup_initial_state()
for i in range(min(len(ck.dt), tscheme.nexp)):
    tscheme.dt[i] = ck.dt[i]
tscheme.set_weights()
dt = tscheme.dt[0].item()
step_time.initialize_dt(dt)
m = {"w": dt_fields.dwdt, "z": dt_fields.dzdt, "p": dt_fields.dpdt, "s": dt_fields.dsdt,
     "b": dt_fields.dbdt, "aj": dt_fields.djdt}
for name, d in ck.derivatives.items():
    for i, arr in enumerate(d["expl"]):
        m[name].expl[:, :, i + 1].copy_(torch.from_numpy(arr).to(dtype=CDTYPE, device=DEVICE))
en = []
for n in range(1, case["cont"] + 1):
    dt = step_time.one_step(n, dt)
    ek = output.get_e_kin_full(fields.w_LMloc, fields.dw_LMloc, fields.z_LMloc)
    em = output.get_e_mag_oc_full(fields.b_LMloc, fields.db_LMloc, fields.aj_LMloc)
    en.append([ek.e_p, ek.e_t, em.e_p, em.e_t])
res = dict(np.load(out + ".npz"))
res["energies_restart"] = np.array(en)
res["w_restart"] = fields.w_LMloc.numpy().copy()
np.savez(out + ".npz", **res)
"""


# Restart files of a conducting, rotating inner core (CNAB2 and BPR353): the
# reference's Fortran checkpoint (MULTISTEP with the inner-core and omega_ic
# histories, DIRK with fields only), its G_1 graph file and its .pt restart file
# after 10 steps, the state they were written from, 10 further steps in memory,
# and 10 steps after each restart: from the Fortran file the way main.run's
# fortran_restart does it (main.py:318-374) and from the .pt file through
# main.load_checkpoint (main.py:226-267), with the Courant cap of
# initialize_dt(dtmax) set by hand (a fresh process would keep it at 0)
IC_FILES_CASES = {
    "filesrot8": dict(cfg=dict(l_max=8, n_r_max=17, n_cheb_max=15, n_cheb_ic_max=7,
                               sigma_ratio=1.0, kbotb=3, nRotIC=1, dtmax=2e-4),
                      dt=2e-4, steps=10, noise=dict(seed=37, amp=1e-2), cont=10),
    "filesbprrot8": dict(cfg=dict(l_max=8, n_r_max=17, n_cheb_max=15, n_cheb_ic_max=7,
                                  sigma_ratio=1.0, kbotb=3, nRotIC=1, dtmax=2e-4,
                                  time_scheme="BPR353"),
                         dt=2e-4, steps=10, noise=dict(seed=41, amp=1e-2), cont=10),
}

IC_FILES_CHILD = r"""
import json, sys
import numpy as np
case = json.loads(sys.argv[1]); out = sys.argv[2]
sys.path.insert(0, sys.argv[3]); sys.path.insert(0, sys.argv[4])
from magic_torch.config import configure
configure(dict(case["cfg"], device="cpu", use_shtns=False))
import torch
torch.set_num_threads(8)
from magic_torch import fields, dt_fields, step_time, output
import magic_torch.main as _m
from magic_torch.init_fields import initialize_fields
from magic_torch.checkpoint_io import write_checkpoint_fortran
from magic_torch.graph_output import write_graph_file
from oracle.params import Params
from oracle import horizontal, radial
from oracle.dynamo import seeded_noise
initialize_fields()
p = Params(case["cfg"])
nz = seeded_noise(p, horizontal.lm_maps(p), radial.radial_functions(p),
                  case["noise"]["seed"], case["noise"]["amp"])
fields.s_LMloc += torch.from_numpy(nz)
step_time.setup_initial_state()
step_time.initialize_dt(case["dt"])
dt = case["dt"]; t = 0.0
for n in range(1, case["steps"] + 1):
    dt = step_time.one_step(n, dt)
    t += dt
write_checkpoint_fortran(out + "_checkpoint.bin", t, case["steps"] + 1)
write_graph_file(out + "_G_1.bin", t, fields.w_LMloc, fields.dw_LMloc, fields.z_LMloc,
                 fields.s_LMloc, fields.p_LMloc, fields.b_LMloc, fields.db_LMloc,
                 fields.aj_LMloc)
_m.save_checkpoint(out + "_restart.pt", case["steps"], t, case["cfg"])
res = dict(time=np.array(t), dt=np.array([dt]), omega_ic=np.array(float(fields.omega_ic)))
for k in ("w", "dw", "z", "p", "s", "b", "db", "aj"):
    res[k] = getattr(fields, k + "_LMloc").numpy().copy()
for k in ("b_ic", "aj_ic"):
    res[k] = getattr(fields, k).numpy().copy()
import magic_torch.time_scheme as ts
res["tscheme_dt"] = ts.tscheme.dt.numpy().copy()
en, om = [], []
for n in range(case["steps"] + 1, case["steps"] + case["cont"] + 1):
    dt = step_time.one_step(n, dt)
    ek = output.get_e_kin_full(fields.w_LMloc, fields.dw_LMloc, fields.z_LMloc)
    em = output.get_e_mag_oc_full(fields.b_LMloc, fields.db_LMloc, fields.aj_LMloc)
    en.append([ek.e_p, ek.e_t, em.e_p, em.e_t])
    om.append(float(fields.omega_ic))
res["energies_cont"] = np.array(en)
res["omega_cont"] = np.array(om)
np.savez(out + ".npz", **res)
"""

IC_RESTART_CHILD = r"""
import json, sys
import numpy as np
case = json.loads(sys.argv[1]); out = sys.argv[2]; kind = sys.argv[4]
sys.path.insert(0, sys.argv[3])
from magic_torch.config import configure
configure(dict(case["cfg"], device="cpu", use_shtns=False))
import torch
torch.set_num_threads(8)
from magic_torch import fields, dt_fields, step_time, output
import magic_torch.main as _m
from magic_torch.time_scheme import tscheme
from magic_torch.precision import CDTYPE, DEVICE
from magic_torch import checkpoint_io
_read = checkpoint_io.read_checkpoint
def _native(path):
    # native-order copies of the big-endian arrays (torch.from_numpy refuses them)
    ck = _read(path)
    nat = lambda a: a.astype(a.dtype.newbyteorder("=")) if isinstance(a, np.ndarray) else a
    for k in ("w", "z", "p", "s", "xi", "b", "aj", "b_ic", "aj_ic", "dt", "r"):
        setattr(ck, k, nat(getattr(ck, k)))
    ck.derivatives = {n: {part: [nat(a) for a in v] for part, v in d.items()}
                      for n, d in ck.derivatives.items()}
    return ck
checkpoint_io.read_checkpoint = _native
if kind == "fortran":
    # main.run's fortran_restart branch (main.py:318-374), TensorBoard aside
    ck = _m.load_fortran_checkpoint(out + "_checkpoint.bin")
    step_time.setup_initial_state()
    for i in range(min(len(ck.dt), tscheme.nexp)):
        tscheme.dt[i] = ck.dt[i]
    tscheme.set_weights()
    dt = tscheme.dt[0].item()
    step_time.initialize_dt(dt)
    m = {"w": dt_fields.dwdt, "z": dt_fields.dzdt, "p": dt_fields.dpdt, "s": dt_fields.dsdt,
         "b": dt_fields.dbdt, "aj": dt_fields.djdt, "b_ic": dt_fields.dbdt_ic,
         "aj_ic": dt_fields.djdt_ic}
    for name, d in ck.derivatives.items():
        for part in ("expl", "impl", "old"):
            for i, arr in enumerate(d[part]):
                getattr(m[name], part)[:, :, i + 1].copy_(
                    torch.from_numpy(arr).to(dtype=CDTYPE, device=DEVICE))
    if ck.domega_ic_dt is not None:
        for part in ("expl", "impl", "old"):
            for i, v in enumerate(ck.domega_ic_dt[part]):
                getattr(dt_fields.domega_ic_dt, part)[i + 1] = v
    start = 0
else:
    start, _ = _m.load_checkpoint(out + "_restart.pt")
    dt = tscheme.dt[0].item()
    # a fresh process keeps the Courant cap at 0.0 (step_time.py:1551), so the
    # reference's first step after this load would take dt = 0: give it the cap
    # initialize_dt(dtmax) sets (this build does the same on a load)
    from magic_torch.pre_calculations import ekScaled
    step_time._dtMax = min(case["cfg"]["dtmax"], tscheme.intfac * ekScaled)
en, om = [], []
for n in range(start + 1, start + case["cont"] + 1):
    dt = step_time.one_step(n, dt)
    ek = output.get_e_kin_full(fields.w_LMloc, fields.dw_LMloc, fields.z_LMloc)
    em = output.get_e_mag_oc_full(fields.b_LMloc, fields.db_LMloc, fields.aj_LMloc)
    en.append([ek.e_p, ek.e_t, em.e_p, em.e_t])
    om.append(float(fields.omega_ic))
res = dict(np.load(out + ".npz"))
res[f"energies_{kind}"] = np.array(en)
res[f"omega_{kind}"] = np.array(om)
res[f"w_{kind}"] = fields.w_LMloc.numpy().copy()
res[f"b_ic_{kind}"] = fields.b_ic.numpy().copy()
np.savez(out + ".npz", **res)
"""


MOVIES = ["Br CMB", "Vr CMB", "Vp CMB", "T CMB", "Br EQ", "Vr EQ", "T EQ", "Bt EQ"]

MOVIE_CHILD = r"""
import json, sys
import numpy as np
case = json.loads(sys.argv[1]); out = sys.argv[2]; movies = json.loads(sys.argv[5])
sys.path.insert(0, sys.argv[3])
from magic_torch.config import configure
configure(dict(case["cfg"], device="cpu", use_shtns=False))
import torch
torch.set_num_threads(8)
from magic_torch import fields
from magic_torch.sht import scal_to_spat, torpol_to_spat
from magic_torch.horizontal_data import dLh
from magic_torch.movie_output import (parse_movie_string, write_movie_header,
                                      write_movie_frame, extract_surface_field,
                                      extract_equat_field)
st = dict(np.load(out + ".npz"))
for k in ("w", "dw", "z", "s", "b", "db", "aj"):
    getattr(fields, k + "_LMloc").copy_(torch.from_numpy(st[k]))
d = dLh.reshape(-1, 1).to(torch.complex128)
# the grid fields of the radial loop (step_time.py:955-1059): velocity, field, entropy
vr, vt, vp = torpol_to_spat(d * fields.w_LMloc, fields.dw_LMloc, fields.z_LMloc)
br, bt, bp = torpol_to_spat(d * fields.b_LMloc, fields.db_LMloc, fields.aj_LMloc)
sc = scal_to_spat(fields.s_LMloc)
t = float(st["time"])
for ms in movies:
    ft, ns, const, fname = parse_movie_string(ms)
    with open(f"{out}_{fname}.bin", "wb") as f:
        write_movie_header(f, ns, 1, const, [ft])
        for k, tk in enumerate((t, 2.0 * t)):
            if ns == 1:
                frame = extract_surface_field(ft, 0, vr, vt, vp, br, bt, bp, sc)
            else:
                frame = extract_equat_field(ft, vr, vt, vp, br, bt, bp, sc)
            write_movie_frame(f, tk, [frame])
"""


def main(names=None):
    env = dict(os.environ, MAGIC_DEVICE="cpu", PYTHONDONTWRITEBYTECODE="1",
               TORCH_COMPILE_DISABLE="1", TORCHDYNAMO_DISABLE="1")
    src = "rvilim/torch-magic src/magic_torch (FP64 CPU)"
    for name, case in CASES.items():
        if names and name not in names:
            continue
        out = os.path.join(HERE, f"traj_{name}.npz")
        subprocess.run([sys.executable, "-c", CHILD, json.dumps(case), out, REF_SRC, REPO],
                       check=True, env=env)
        meta = dict(case, source=src + ": step_time.one_step, dt chained")
        with open(os.path.join(HERE, f"traj_{name}.json"), "w") as f:
            json.dump(meta, f, indent=1)
        print("wrote", out)
    if not names or "files16" in names:
        out = os.path.join(HERE, "files16")
        subprocess.run([sys.executable, "-c", FILES_CHILD, json.dumps(FILES_CASE), out, REF_SRC,
                        REPO], check=True, env=env)
        subprocess.run([sys.executable, "-c", RESTART_CHILD, json.dumps(FILES_CASE), out,
                        REF_SRC], check=True, env=env)
        meta = dict(FILES_CASE, source=src + ": checkpoint_io.write_checkpoint_fortran, "
                    "graph_output.write_graph_file, main.run fortran_restart sequence")
        with open(out + ".json", "w") as f:
            json.dump(meta, f, indent=1)
        print("wrote", out)
    if not names or "movies16" in names:
        out = os.path.join(HERE, "files16")
        subprocess.run([sys.executable, "-c", MOVIE_CHILD, json.dumps(FILES_CASE), out, REF_SRC,
                        REPO, json.dumps(MOVIES)], check=True, env=env)
        print("wrote movie files of", out)
    for name, case in IC_FILES_CASES.items():
        if names and name not in names:
            continue
        out = os.path.join(HERE, name)
        subprocess.run([sys.executable, "-c", IC_FILES_CHILD, json.dumps(case), out, REF_SRC,
                        REPO], check=True, env=env)
        for kind in ("fortran", "pt"):
            subprocess.run([sys.executable, "-c", IC_RESTART_CHILD, json.dumps(case), out,
                            REF_SRC, kind], check=True, env=env)
        meta = dict(case, source=src + ": checkpoint_io.write_checkpoint_fortran, "
                    "graph_output.write_graph_file, main.save_checkpoint, main.run "
                    "fortran_restart sequence, main.load_checkpoint")
        with open(out + ".json", "w") as f:
            json.dump(meta, f, indent=1)
        print("wrote", out)
    for name, case in FD_CASES.items():
        if names and name not in names:
            continue
        out = os.path.join(HERE, f"{name}.npz")
        subprocess.run([sys.executable, "-c", FD_CHILD, json.dumps(case), out, REF_SRC],
                       check=True, env=env)
        meta = dict(case, source=src + ": build_s/z/w_matrices, FD branch")
        with open(os.path.join(HERE, f"{name}.json"), "w") as f:
            json.dump(meta, f, indent=1)
        print("wrote", out)


if __name__ == "__main__":
    main(sys.argv[1:] or None)
