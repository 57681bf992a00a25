"""Anelastic (polytropic) Chebyshev step pieces (oracle restatement, FP64 numpy).

Test infrastructure: restates the reference's anelastic branch for the
checker; the product path never imports it.

* ``background`` -- the reference state on the Chebyshev grid
  (``radial_functions.py:71-167``): polytropic temp0, rho0, beta, dbeta,
  dLtemp0, orho1, otemp1, alpha0 = 1/temp0, ogrun = polind, gravity
  g0 + g1 r/r_cmb + g2 r_cmb^2/r^2, ViscHeatFac = DissNb pr/ra and
  OhmLossFac = ViscHeatFac/(ek prmag^2); constant transport properties
  (visc = kappa = 1).  Boussinesq runs get the trivial profiles.
* ``build_s`` / ``build_z`` / ``build_wp`` / ``build_p0`` -- the anelastic
  operator matrices (``update_s.py:66-188``, ``update_z.py:89-240``,
  ``update_wp.py:197-368`` with row and column preconditioning,
  ``update_wp.py:80-170`` with the Chebyshev integral condition on the l=0
  pressure when ViscHeatFac ThExpNb != 0); stored as preconditioned inverses
  like the Boussinesq ones (the reference LU-solves the (w, p) systems: the
  same solution).
* ``ps_cond`` -- the anelastic conduction state (``init_fields.py:190-305``).
* ``get_nl`` -- non-curl advection, Lorentz force, v x B, v s and viscous
  (and Ohmic) heating on the grid (``get_nl_anel.py:41-136``).
* ``am_correct`` -- the angular-momentum corrections of z(l=1, m=0) and
  z(l=1, m=1) (``update_z.py:301-331,450-475``).

Pinned by ``tests/test_oracle.py`` against reference trajectories of the
anelastic step (``traj_anel*`` written by ``tests/golden/make_golden.py``).
"""

from __future__ import annotations

import math

import numpy as np

from . import fd
from .radial import costf


def background(p, rad) -> dict:
    """Radial profiles of the reference state on the run's grid."""
    r = rad["r"]
    n = r.shape[0]
    ones, zeros = np.ones(n), np.zeros(n)
    if p.l_anel:
        prof = fd.polytropic(r, rad["r_cmb"], rad["r_icb"], p.radratio, p.strat, p.polind,
                             p.g0, p.g1, p.g2)
        t0 = prof["temp0"]
        vhf = prof["DissNb"] * p.pr / p.ra
        prof.update(otemp1=1.0 / t0, alpha0=1.0 / t0, ogrun=np.full(n, p.polind),
                    ViscHeatFac=vhf, OhmLossFac=vhf / (p.ek * p.prmag ** 2) if p.l_mag else 0.0)
    else:
        prof = dict(rgrav=rad["rgrav"], temp0=ones, rho0=ones, beta=zeros, dbeta=zeros,
                    ddbeta=zeros, dLtemp0=zeros, orho1=ones, visc=ones, dLvisc=zeros,
                    ddLvisc=zeros, kappa=ones, dLkappa=zeros, otemp1=ones, alpha0=ones,
                    ogrun=zeros, DissNb=0.0, ViscHeatFac=0.0, OhmLossFac=0.0)
    prof.update(or1=rad["or1"], or2=rad["or2"], or3=rad["or3"], or4=rad["or4"], ThExpNb=1.0)
    return prof


def _bfac(dat, cols):
    for c in cols:
        dat[:, c] *= 0.5
    return dat


def _inv_rows(dat):
    fac = 1.0 / np.abs(dat).max(axis=1)
    return np.linalg.inv(fac[:, None] * dat) * fac[None, :]


def build_s(p, rad, mats, prof, w0):
    """sMat per l with the anelastic first-derivative and kappa terms."""
    R, D, D2, _ = mats
    N, rn = p.n_r_max, rad["rnorm"]
    d1 = (prof["beta"] + prof["dLtemp0"] + 2.0 * prof["or1"] + prof["dLkappa"])[:, None]
    kap = prof["kappa"][:, None]
    or2 = prof["or2"][:, None]
    out = np.zeros((p.l_max + 1, N, N))
    for l in range(p.l_max + 1):
        dL = float(l * (l + 1))
        dat = rn * (R - w0 * p.opr * kap * (D2 + d1 * D - dL * or2 * R))
        dat[0] = rn * R[0]
        dat[N - 1] = rn * R[N - 1]
        if p.n_cheb_max < N:
            dat[(0, N - 1), p.n_cheb_max:] = 0.0
        out[l] = _inv_rows(_bfac(dat, (0, N - 1)))
    return out


def build_z(p, rad, mats, prof, w0):
    """zMat per l >= 1 with beta, dbeta (stress-free rows carry beta)."""
    R, D, D2, _ = mats
    N, rn = p.n_r_max, rad["rnorm"]
    c = lambda k: prof[k][:, None]
    or1, or2, beta, dbeta, dLv, visc = c("or1"), c("or2"), c("beta"), c("dbeta"), c("dLvisc"), c("visc")
    o1, b = prof["or1"], prof["beta"]
    out = np.zeros((p.l_max + 1, N, N))
    for l in range(1, p.l_max + 1):
        dL = float(l * (l + 1))
        dat = rn * dL * or2 * (R - w0 * visc * (
            D2 + (dLv - beta) * D
            - (dLv * beta + 2.0 * dLv * or1 + dL * or2 + dbeta + 2.0 * beta * or1) * R))
        dat[0] = rn * (D[0] - (2.0 * o1[0] + b[0]) * R[0]) if p.ktopv == 1 else rn * R[0]
        dat[N - 1] = (rn * (D[N - 1] - (2.0 * o1[N - 1] + b[N - 1]) * R[N - 1]) if p.kbotv == 1
                      else rn * R[N - 1])
        if p.n_cheb_max < N:
            dat[(0, N - 1), p.n_cheb_max:] = 0.0
        out[l] = _inv_rows(_bfac(dat, (0, N - 1)))
    return out


def build_wp(p, rad, mats, prof, w0):
    """Coupled (w, p) matrices per l >= 1, anelastic terms, row and column
    preconditioning; returns the (l_max+1, 2N, 2N) inverses."""
    R, D, D2, D3 = mats
    N, rn = p.n_r_max, rad["rnorm"]
    b = slice(1, N - 1)
    c = lambda k: prof[k][b, None]
    o1, o2, be, dbe, dLv, vi = c("or1"), c("or2"), c("beta"), c("dbeta"), c("dLvisc"), c("visc")
    or1, bt = prof["or1"], prof["beta"]
    out = np.zeros((p.l_max + 1, 2 * N, 2 * N))
    for l in range(1, p.l_max + 1):
        dL = float(l * (l + 1))
        dat = np.zeros((2 * N, 2 * N))
        dat[b, :N] = rn * dL * o2 * (R[b] - w0 * vi * (
            D2[b] + (2.0 * dLv - be / 3.0) * D[b]
            - (dL * o2 + 4.0 / 3.0 * (dLv * be + (3.0 * dLv + be) * o1 + dbe)) * R[b]))
        dat[b, N:] = rn * w0 * (D[b] - be * R[b])
        dat[N + 1:2 * N - 1, :N] = rn * dL * o2 * (-D[b] - w0 * vi * (
            -D3[b] + (be - dLv) * D2[b]
            + (dL * o2 + dbe + dLv * be + 2.0 * (dLv + be) * o1) * D[b]
            - dL * o2 * (2.0 * o1 + dLv + 2.0 / 3.0 * be) * R[b]))
        dat[N + 1:2 * N - 1, N:] = -rn * w0 * dL * o2 * R[b]
        dat[0, :N] = rn * R[0]
        dat[N - 1, :N] = rn * R[N - 1]
        dat[N, :N] = (rn * (D2[0] - (2.0 * or1[0] + bt[0]) * D[0]) if p.ktopv == 1
                      else rn * D[0])
        dat[2 * N - 1, :N] = (rn * (D2[N - 1] - (2.0 * or1[N - 1] + bt[N - 1]) * D[N - 1])
                              if p.kbotv == 1 else rn * D[N - 1])
        if p.n_cheb_max < N:
            for row in (0, N - 1, N, 2 * N - 1):
                dat[row, p.n_cheb_max:N] = 0.0
                dat[row, N + p.n_cheb_max:] = 0.0
        _bfac(dat, (0, N - 1, N, 2 * N - 1))
        rm = np.abs(dat).max(axis=1)
        rm[rm == 0] = 1.0
        fr = 1.0 / rm
        dat = fr[:, None] * dat
        cm = np.abs(dat).max(axis=0)
        cm[cm == 0] = 1.0
        fc = 1.0 / cm
        dat = dat * fc[None, :]
        out[l] = np.linalg.inv(dat) * fc[:, None] * fr[None, :]
    return out


def cheb_integ_kernel(n: int, rnorm: float) -> np.ndarray:
    """K[j, k] = (1/(1-(k-j)^2) + 1/(1-(k+j)^2)) rnorm/2 for j + k even, else 0:
    the integral of T_j T_k-weighted coefficients (``init_fields.py:84-106``)."""
    j = np.arange(n, dtype=np.float64)[:, None]
    k = np.arange(n, dtype=np.float64)[None, :]
    d1, d2 = 1.0 - (k - j) ** 2, 1.0 - (k + j) ** 2
    with np.errstate(divide="ignore"):
        s1 = np.where(d1 != 0, 1.0 / np.where(d1 != 0, d1, 1.0), 0.0)
        s2 = np.where(d2 != 0, 1.0 / np.where(d2 != 0, d2, 1.0), 0.0)
    return ((j + k) % 2 == 0) * (s1 + s2) * 0.5 * rnorm


def _integ_row(rad, prof_phys, n):
    """Chebyshev integral row of a physical profile (the bfac-scaled spectral
    coefficients through the integration kernel)."""
    rn = rad["rnorm"]
    work = costf(prof_phys) * rn
    work[0] *= 0.5
    work[n - 1] *= 0.5
    return cheb_integ_kernel(n, rn) @ work


def p0_integral_bc(p, prof) -> bool:
    return p.l_anel and prof["ViscHeatFac"] * prof["ThExpNb"] != 0.0


def build_p0(p, rad, mats, prof):
    """l=0 pressure matrix inverse: rows 1.. rnorm (D - beta R), row 0 the
    integral condition (or p = 0 at the CMB)."""
    R, D, _, _ = mats
    N, rn = p.n_r_max, rad["rnorm"]
    r = rad["r"]
    dat = np.zeros((N, N))
    dat[1:] = rn * (D[1:] - prof["beta"][1:, None] * R[1:])
    if p0_integral_bc(p, prof):
        dat[0] = _integ_row(rad, prof["ThExpNb"] * prof["ViscHeatFac"] * prof["ogrun"]
                            * prof["alpha0"] * r * r, N)
    else:
        dat[0] = rn * R[0]
    if p.n_cheb_max < N:
        dat[0, p.n_cheb_max:] = 0.0
    return np.linalg.inv(_bfac(dat, (0, N - 1)))


def p0_rhs_weights(p, rad, prof, q):
    """Weights u with p0_rhs[0] = u @ Re s(l=0) (``update_wp.py:431-436``)."""
    r = rad["r"]
    return prof["ThExpNb"] * prof["alpha0"] * prof["temp0"] * prof["rho0"] * r * r * q


def ps_cond(p, rad, mats, prof):
    """Anelastic conduction state (s0, p0) on the grid (``init_fields.py:190-305``)."""
    from .algebra import prepare_mat, solve_mat
    R, D, D2, _ = mats
    N, rn = p.n_r_max, rad["rnorm"]
    r = rad["r"]
    M = np.zeros((2 * N, 2 * N))
    d1 = (prof["beta"] + prof["dLtemp0"] + 2.0 * prof["or1"] + prof["dLkappa"])[:, None]
    M[:N, :N] = rn * p.opr * prof["kappa"][:, None] * (D2 + d1 * D)
    M[N:, :N] = -rn * (prof["rho0"] * p.BuoFac * prof["rgrav"])[:, None] * R
    M[N:, N:] = rn * (D - prof["beta"][:, None] * R)
    M[0, :N] = rn * R[0]; M[0, N:] = 0.0
    M[N - 1, :N] = rn * R[N - 1]; M[N - 1, N:] = 0.0
    if prof["ViscHeatFac"] * prof["ThExpNb"] != 0.0:
        M[N, N:] = _integ_row(rad, prof["ThExpNb"] * prof["ViscHeatFac"] * prof["ogrun"]
                              * prof["alpha0"] * r * r, N)
        M[N, :N] = _integ_row(rad, -prof["ThExpNb"] * prof["alpha0"] * prof["temp0"]
                              * prof["rho0"] * r * r, N)
    else:
        M[N, :N] = 0.0
        M[N, N:] = rn * R[0]
    for col in (0, N - 1, N, 2 * N - 1):
        M[:, col] *= 0.5
    fac = 1.0 / np.abs(M).max(axis=1)
    M = M * fac[:, None]
    a, ip, info = prepare_mat(M)
    assert info == 0, f"Singular ps0Mat (anelastic), info={info}"
    rhs = np.zeros(2 * N)
    rhs[N - 1] = math.sqrt(4.0 * math.pi)
    x = solve_mat(a, ip, fac * rhs)
    return costf(x[:N]), costf(x[N:])


def get_nl(p, prof, grid, vr, vt, vp, dvrdr, dvtdr, dvpdr, dvrdt, dvrdp, dvtdp, dvpdp, cvr,
           sc, br, bt, bp, cbr, cbt, cbp):
    """Grid products of the anelastic step (``get_nl_anel.py:41-136``): inputs
    (n_r, n_theta, n_phi); magnetic inputs may be zeros."""
    c3 = lambda k: prof[k][:, None, None]
    or1, or2, or4, beta, orho1 = c3("or1"), c3("or2"), c3("or4"), c3("beta"), c3("orho1")
    otemp1, visc = c3("otemp1"), c3("visc")
    r = 1.0 / or1
    ost2 = grid["O_sin_theta_E2_grid"][None, :, None]
    cosn = (grid["cos_grid"] * grid["O_sin_theta_E2_grid"])[None, :, None]
    LF = p.LFfac
    LFr = LF * ost2 * (cbt * bp - cbp * bt)
    LFt = LF * or4 * (cbp * br - cbr * bp)
    LFp = LF * or4 * (cbr * bt - cbt * br)
    Advr = -or2 * orho1 * (vr * (dvrdr - (2.0 * or1 + beta) * vr)
                           + ost2 * (vt * (dvrdt - r * vt) + vp * (dvrdp - r * vp))) + LFr
    Advt = or4 * orho1 * (-vr * (dvtdr - beta * vt) + vt * (cosn * vt + dvpdp + dvrdr)
                          + vp * (cosn * vp - dvtdp)) + LFt
    Advp = or4 * orho1 * (-vr * (dvpdr - beta * vp) - vt * (dvtdp + cvr) - vp * dvpdp) + LFp
    VSr = vr * sc
    VSt = or2 * vt * sc
    VSp = or2 * vp * sc
    VxBr = orho1 * ost2 * (vt * bp - vp * bt)
    VxBt = orho1 * or4 * (vp * br - vr * bp)
    VxBp = orho1 * or4 * (vr * bt - vt * br)
    prefac = prof["ViscHeatFac"] * or4 * orho1 * otemp1 * visc
    br_ = beta * r
    e_rr = dvrdr - (2.0 * or1 + beta) * vr
    e_tt = cosn * vt + dvpdp + dvrdr - or1 * vr
    e_pp = dvpdp + cosn * vt + or1 * vr
    e_tp = 2.0 * dvtdp + cvr - 2.0 * cosn * vp
    e_rt = r * dvtdr - (2.0 + br_) * vt + or1 * dvrdt
    e_rp = r * dvpdr - (2.0 + br_) * vp + or1 * dvrdp
    heat = prefac * (2.0 * e_rr * e_rr + 2.0 * e_tt * e_tt + 2.0 * e_pp * e_pp + e_tp * e_tp
                     + ost2 * (e_rt * e_rt + e_rp * e_rp) - 2.0 / 3.0 * (beta * vr) ** 2)
    if p.l_mag and prof["OhmLossFac"] != 0.0:
        heat = heat + prof["OhmLossFac"] * or2 * otemp1 * (
            or2 * cbr * cbr + ost2 * (cbt * cbt + cbp * cbp))
    return Advr, Advt, Advp, VSr, VSt, VSp, VxBr, VxBt, VxBp, heat


def am_constants(p, rad, prof, q):
    """Profiles and moments of the angular-momentum corrections
    (``update_z.py:57-66``, ``pre_calculations.py:60-105``)."""
    r = rad["r"]
    rho0, beta, dbeta = prof["rho0"], prof["beta"], prof["dbeta"]
    r2 = r * r
    return dict(r2q=r2 * q, z=rho0 * r2, dz=rho0 * (2.0 * r + r2 * beta),
                d2z=rho0 * (2.0 + 4.0 * beta * r + dbeta * r2 + beta * beta * r2),
                c_moi_oc=8.0 / 3.0 * math.pi * float((r ** 4 * rho0) @ q),
                y10=0.5 * math.sqrt(3.0 / math.pi), y11=0.5 * math.sqrt(1.5 / math.pi))


def am_correct(p, am, z, dz, d2z, l1m0, l1m1, omega_ic=0.0, c_moi_ic=0.0):
    """Axial (l_correct_AMz) and equatorial (l_correct_AMe) angular-momentum
    corrections of z, dz, d2z in place (``update_z.py:450-475``; AMstart = 0,
    no mantle rotation)."""
    fac = 8.0 / 3.0 * math.pi
    if p.l_correct_AMz:
        am_z = fac * am["y10"] * float(z[l1m0].real @ am["r2q"]) + c_moi_ic * omega_ic
        corr = am_z / (am["c_moi_oc"] * am["y10"])
        z[l1m0] -= am["z"] * corr
        dz[l1m0] -= am["dz"] * corr
        d2z[l1m0] -= am["d2z"] * corr
    if p.l_correct_AMe and l1m1 >= 0:
        am_x = 2.0 * fac * am["y11"] * float(z[l1m1].real @ am["r2q"])
        am_y = -2.0 * fac * am["y11"] * float(z[l1m1].imag @ am["r2q"])
        corr = complex(am_x, -am_y) / (2.0 * am["y11"] * am["c_moi_oc"])
        z[l1m1] -= am["z"] * corr
        dz[l1m1] -= am["dz"] * corr
        d2z[l1m1] -= am["d2z"] * corr
