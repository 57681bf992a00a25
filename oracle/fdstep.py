"""Finite-difference (double-curl) time step pieces (oracle restatement, FP64 numpy).

Test infrastructure: restates the reference's finite-difference branches for
the checker; the product path never imports it.

With ``radial_scheme = "FD"`` the reference runs the double-curl form of the
poloidal equation (``params.py:143``): pressure is eliminated for l >= 1,
w obeys a fourth-order equation and only the l = 0 pressure is solved.  The
grid is the stretched finite-difference grid, rMat = I, rnorm = boundary_fac
= 1, no Chebyshev transform, and radial integrals are Simpson's rule.

* ``build`` -- per-degree inverses of the S, Z, double-curl W, insulating B and
  J operators and the l = 0 pressure matrix (``update_s.py:121-176``,
  ``update_z.py:89-240``, ``update_w_doublecurl.py:67-307``,
  ``update_b.py:160-262``, ``update_wp.py:95-170``: last row first-order
  backward difference cut to the Fortran band, Dirichlet row 0).  The
  reference solves through dense inverses (N <= 1024) or band LU: the same
  solution up to rounding.
* ``ps_cond`` -- the conduction state (``init_fields.py:108-182`` Boussinesq,
  ``init_fields.py:183-311`` anelastic with the trapezoidal integral row).
* ``td_w`` -- the double-curl explicit poloidal term (``get_td.py:91-147``
  + ``update_wp.py:371-384``).
* ``w_old`` / ``w_impl`` -- old state and implicit terms of w
  (``update_w_doublecurl.py:309-393``).

Pinned by ``tests/test_oracle_fdstep.py`` against reference trajectories of
the finite-difference step (``traj_fd*`` written by
``tests/golden/make_golden.py``) and by the operator fixtures ``fd4_ob*``.
"""

from __future__ import annotations

import math

import numpy as np

from . import fd


def _inv_rows(dat):
    fac = 1.0 / np.abs(dat).max(axis=1)
    return np.linalg.inv(fac[:, None] * dat) * fac[None, :]


def build(p, rad, mats, prof, w0):
    """Inverses by degree: s, z, w (l >= 1; zero at l = 0), b, j (magnetic runs)
    and p0, each acting on grid values and returning grid values."""
    N, lp1 = p.n_r_max, p.l_max + 1
    R, D1, D2 = mats[0], mats[1], mats[2]
    inv = dict(s=np.zeros((lp1, N, N)), z=np.zeros((lp1, N, N)), w=np.zeros((lp1, N, N)))
    for l in range(lp1):
        dat, fac = fd.s_operator(l, w0, p.opr, mats, prof)
        inv["s"][l] = np.linalg.inv(dat) * fac[None, :]
        if l == 0:
            continue
        dat, fac = fd.z_operator(l, w0, mats, prof, p.ktopv, p.kbotv)
        inv["z"][l] = np.linalg.inv(dat) * fac[None, :]
        dat, fr, fc = fd.w_operator(l, w0, mats, prof, p.ktopv, p.kbotv)
        inv["w"][l] = fc[:, None] * np.linalg.inv(dat) * fr[None, :]
    if p.l_mag:
        or1, or2 = prof["or1"], prof["or2"]
        inv["b"], inv["j"] = np.zeros((lp1, N, N)), np.zeros((lp1, N, N))
        for l in range(1, lp1):
            dL = float(l * (l + 1))
            bulk = dL * or2[:, None] * (R - w0 * p.opm * (D2 - dL * or2[:, None] * R))
            db = bulk.copy()
            db[0] = D1[0] + l * or1[0] * R[0]
            db[N - 1] = D1[N - 1] - (l + 1) * or1[N - 1] * R[N - 1]
            dj = bulk.copy()
            dj[0] = R[0]
            dj[N - 1] = R[N - 1]
            inv["b"][l] = _inv_rows(db)
            inv["j"][l] = _inv_rows(dj)
    r, beta = rad["r"], prof["beta"]
    d = np.zeros((N, N))
    d[1:] = D1[1:] - beta[1:, None] * R[1:]
    delr = r[N - 1] - r[N - 2]
    d[N - 1, N - 1] = 1.0 / delr - beta[N - 1]
    d[N - 1, N - 2] = -1.0 / delr
    d[N - 1, :N - 1 - p.fd_order // 2] = 0.0
    d[0] = R[0]
    inv["p0"] = np.linalg.inv(d)
    return inv


def ps_cond(p, rad, mats, prof):
    """(s0, p0) of the conduction state on the grid: the 2N system of entropy
    diffusion and hydrostatic balance, s = 0 on the outer and sqrt(4 pi) on the
    inner boundary; row N fixes p(r_cmb) = 0, or, in anelastic runs with
    ViscHeatFac ThExpNb != 0, the trapezoidal integral of the pressure and
    entropy weights."""
    N = p.n_r_max
    R, D1, D2 = mats[0], mats[1], mats[2]
    r = rad["r"]
    M = np.zeros((2 * N, 2 * N))
    d1 = (prof["beta"] + prof["dLtemp0"] + 2.0 * prof["or1"] + prof["dLkappa"])[:, None]
    M[:N, :N] = p.opr * prof["kappa"][:, None] * (D2 + d1 * D1)
    M[N:, :N] = -(prof["rho0"] * p.BuoFac * prof["rgrav"])[:, None] * R
    M[N:, N:] = D1 - prof["beta"][:, None] * R
    M[0, :N] = R[0]; M[0, N:] = 0.0
    M[N - 1, :N] = R[N - 1]; M[N - 1, N:] = 0.0
    vtf = prof["ViscHeatFac"] * prof["ThExpNb"]
    if p.l_anel and vtf != 0.0:
        wp = vtf * prof["ogrun"] * prof["alpha0"] * r * r
        ws = -prof["ThExpNb"] * prof["alpha0"] * prof["temp0"] * prof["rho0"] * r * r
        M[N] = 0.0
        for i in range(1, N - 1):
            h = r[i + 1] - r[i - 1]
            M[N, i] = 0.5 * ws[i] * h
            M[N, N + i] = 0.5 * wp[i] * h
        M[N, 0] = 0.5 * ws[0] * (r[1] - r[0])
        M[N, N] = 0.5 * wp[0] * (r[1] - r[0])
        M[N, N - 1] = 0.5 * ws[N - 1] * (r[N - 1] - r[N - 2])
        M[N, 2 * N - 1] = 0.5 * wp[N - 1] * (r[N - 1] - r[N - 2])
    else:
        M[N, :N] = 0.0
        M[N, N:] = R[0]
    fac = 1.0 / np.abs(M).max(axis=1)
    rhs = np.zeros(2 * N)
    rhs[N - 1] = math.sqrt(4.0 * math.pi)
    sol = np.linalg.solve(fac[:, None] * M, fac * rhs)
    return sol[:N], sol[N:]


def td_w(run, AdvrLM, AdvtLM):
    """Double-curl explicit poloidal term: dLh or4 orho1 Advr plus the Coriolis
    term 2 CorFac or2 orho1 [dPhi (-w'' + beta w' + (beta/r + 1/r^2) dLh w)
    + dTheta3A (z'_A - beta z_A) + dTheta3S (z'_S - beta z_S) + (dTheta4A z_A
    - dTheta4S z_S)/r]; l = 0 keeps the standard form or2 Advr + 2 CorFac/r
    dTheta2A z(1,0); zero on the boundaries; then + or2 d/dr of
    dVxVh = -orho1 r^2 dLh Advt (zero on the boundaries)."""
    p, rad, A, maps, pr = run.prm, run.rad, run.lmarr, run.maps, run.prof
    CorFac = p.CorFac
    or1, or2, or4 = rad["or1"][None], rad["or2"][None], rad["or4"][None]
    r2 = (rad["r"] ** 2)[None]
    orho1, beta = pr["orho1"][None], pr["beta"][None]
    dLh, dPhi = A["dLh"][:, None], A["dPhi"][:, None]
    lmA = np.clip(maps["lm2lmA"], 0, None)
    maskA = (maps["lm2lmA"] >= 0)[:, None].astype(np.float64)
    lmS = maps["lm2lmS"]
    gA = lambda f: f[lmA] * maskA
    gS = lambda f: f[lmS]
    c = {k: A[k][:, None] for k in ("dTheta2A", "dTheta3A", "dTheta3S", "dTheta4A", "dTheta4S")}
    w, dw, ddw, z, dz = run.w, run.dw, run.ddw, run.z, run.dz
    zA, zS = gA(z), gS(z)
    dwdt = dLh * or4 * orho1 * AdvrLM
    dVxVh = -orho1 * r2 * dLh * AdvtLM
    dwdt = dwdt + 2.0 * CorFac * or2 * orho1 * (
        dPhi * (-ddw + beta * dw + (beta * or1 + or2) * dLh * w)
        + c["dTheta3A"] * (gA(dz) - beta * zA) + c["dTheta3S"] * (gS(dz) - beta * zS)
        + or1 * (c["dTheta4A"] * zA - c["dTheta4S"] * zS))
    l0 = run.lm_l0
    dwdt[l0] = or2[0] * AdvrLM[l0] + 2.0 * CorFac * or1[0] * c["dTheta2A"][l0, 0] * zA[l0]
    for f in (dwdt, dVxVh):
        f[:, 0] = 0.0
        f[:, -1] = 0.0
    return dwdt + or2 * (dVxVh @ run.D1.T)


def w_old(run, w, dw, ddw):
    """dLh or2 (-orho1 (w'' - beta w' - dLh or2 w)), zero on the boundaries."""
    rad, pr = run.rad, run.prof
    dLh = run.lmarr["dLh"][:, None]
    or2, orho1, beta = rad["or2"][None], pr["orho1"][None], pr["beta"][None]
    out = dLh * or2 * (-orho1 * (ddw - beta * dw - dLh * or2 * w))
    out[:, 0] = 0.0
    out[:, -1] = 0.0
    return out


def w_impl(run, w, dw, ddw, dddw, ddddw, s):
    """-dLh or2 visc orho1 (w'''' + c3 w''' + c2 w'' + c1 w' + c0 w) + BuoFac dLh
    or2 rgrav s, with the coefficients of the double-curl viscous operator;
    zero on the boundaries."""
    p, rad, pr = run.prm, run.rad, run.prof
    dLh = run.lmarr["dLh"][:, None]
    g = lambda k: pr[k][None]
    or1, or2 = rad["or1"][None], rad["or2"][None]
    be, dbe, ddbe = g("beta"), g("dbeta"), g("ddbeta")
    dv, ddv = g("dLvisc"), g("ddLvisc")
    c3 = 2.0 * (dv - be)
    c2 = (ddv - 2.0 * dbe + dv ** 2 + be ** 2 - 3.0 * dv * be - 2.0 * or1 * (dv + be)
          - 2.0 * dLh * or2)
    c1 = (-ddbe - dbe * (2.0 * dv - be + 2.0 * or1) - ddv * (be + 2.0 * or1)
          + be ** 2 * (dv + 2.0 * or1) - be * (dv ** 2 - 2.0 * or2)
          - 2.0 * dv * or1 * (dv - or1) + 2.0 * (2.0 * or1 + be - dv) * dLh * or2)
    c0 = dLh * or2 * (2.0 * dbe + ddv + dv ** 2 - (2.0 / 3.0) * be ** 2 + dv * be
                      + 2.0 * or1 * (2.0 * dv - be - 3.0 * or1) + dLh * or2)
    dif = -dLh * or2 * g("visc") * g("orho1") * (ddddw + c3 * dddw + c2 * ddw + c1 * dw + c0 * w)
    out = dif + p.BuoFac * dLh * or2 * g("rgrav") * s
    out[:, 0] = 0.0
    out[:, -1] = 0.0
    return out
