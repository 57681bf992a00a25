"""Anelastic (polytropic) reference state and operators of the device step.

With ``strat > 0`` the reference runs the anelastic step
(``step_time.py:1215-1475``): a polytropic background of index ``polind``
(``radial_functions.py:71-167``), non-curl advection with viscous (and Ohmic)
heating in grid space (``get_nl_anel.py:41-136``), 1/rho0-weighted Courant
limits and kinetic energies, and operators with the background's density
scale height beta = d ln rho0 / dr (``update_s.py:66-188``,
``update_z.py:89-240``, ``update_wp.py:80-368``).  Transport properties are
constant (visc = kappa = lambda = 1), the entropy gradient of the background
zero.

The S, Z and (w, p) collocation matrices are affine in the implicit weight:
``affine_operators`` assembles A0 and A1 once per run on the host, and every
rebuild forms A0 + w A1 on the device before the batched inverses (the (w, p)
systems with the reference's row and column scaling; the reference LU-solves
them: the same solution) and the folds with the DCT-I and the radial
derivatives.  The l = 0 pressure (its integral condition when
ViscHeatFac ThExpNb != 0, ``update_wp.py:100-117``, ``init_fields.py:84-106``)
does not depend on the weight: its inverse is folded into a fixed block of the
(w, p) row operator once (``p0_block``).  ``s_inverse`` / ``z_inverse`` /
``wp_inverse`` / ``p0_inverse`` are the host inverses (conduction state,
checks).
"""
from __future__ import annotations

import math

import numpy as np


def background(p, rad: dict) -> dict:
    """Profiles (N,) of the polytropic reference state on the radial grid,
    and its scalar factors."""
    r, rc, ri = rad["r"], rad["r_cmb"], rad["r_icb"]
    n = r.shape[0]
    or1 = 1.0 / r
    or2, or3 = or1 * or1, or1 ** 3
    g0, g1, g2, pi = p.g0, p.g1, p.g2, p.polind
    rgrav = g0 + g1 * r / rc + g2 * rc * rc * or2
    diss = ((math.exp(p.strat / pi) - 1.0) / (rc - ri)
            / (g0 + 0.5 * g1 * (1.0 + p.radratio) + g2 / p.radratio))
    temp0 = -diss * (g0 * r + 0.5 * g1 * r * r / rc - g2 * rc * rc * or1) + 1.0 + diss * rc * (
        g0 + 0.5 * g1 - g2)
    dgrav = g1 / rc - 2.0 * g2 * rc * rc * or3
    beta = -pi * diss * rgrav / temp0
    dbeta = -pi * diss / temp0 ** 2 * (dgrav * temp0 + diss * rgrav ** 2)
    ddbeta = -pi * diss / temp0 ** 3 * (
        temp0 * (6.0 * g2 * rc * rc * or2 * or2 * temp0 + diss * rgrav * dgrav)
        + 2.0 * diss * rgrav * (diss * rgrav ** 2 + dgrav * temp0))
    rho0 = temp0 ** pi
    vhf = diss * p.pr / p.ra
    return dict(rgrav=rgrav, temp0=temp0, rho0=rho0, orho1=1.0 / rho0, otemp1=1.0 / temp0,
                alpha0=1.0 / temp0, ogrun=np.full(n, pi), beta=beta, dbeta=dbeta,
                ddbeta=ddbeta, dLtemp0=-diss * rgrav / temp0, visc=np.ones(n),
                dLvisc=np.zeros(n), ddLvisc=np.zeros(n),
                kappa=np.ones(n), dLkappa=np.zeros(n), DissNb=diss, ViscHeatFac=vhf,
                OhmLossFac=vhf / (p.ek * p.prmag ** 2) if p.l_mag else 0.0, ThExpNb=1.0)


def _rows_inv(a):
    """Inverse of row-scaled systems (..., n, n) with the scaling folded in."""
    f = 1.0 / np.abs(a).max(axis=-1)
    return np.linalg.inv(f[..., :, None] * a) * f[..., None, :]


def _halve(a, cols):
    for c in cols:
        a[..., :, c] *= 0.5
    return a


def s_inverse(p, rad, prof, w0):
    return _rows_inv(_s_matrix(p, rad, prof, w0))


def _s_matrix(p, rad, prof, w0):
    R, D, D2 = rad["T"], rad["T1"], rad["T2"]
    N, rn, lp1 = p.n_r_max, rad["rnorm"], p.l_max + 1
    dL = (np.arange(lp1) * (np.arange(lp1) + 1.0))[:, None, None]
    d1 = (prof["beta"] + prof["dLtemp0"] + 2.0 / rad["r"] + prof["dLkappa"])[:, None]
    a = rn * (R - w0 * p.opr * prof["kappa"][:, None] * (D2 + d1 * D - dL * rad["or2"][:, None] * R))
    a[:, 0] = rn * R[0]
    a[:, N - 1] = rn * R[N - 1]
    if p.n_cheb_max < N:
        a[:, (0, N - 1), p.n_cheb_max:] = 0.0
    return _halve(a, (0, N - 1))


def z_inverse(p, rad, prof, w0):
    out = _rows_inv(_z_matrix(p, rad, prof, w0))
    out[0] = 0.0
    return out


def _z_matrix(p, rad, prof, w0):
    R, D, D2 = rad["T"], rad["T1"], rad["T2"]
    N, rn, lp1 = p.n_r_max, rad["rnorm"], p.l_max + 1
    dL = (np.arange(lp1) * (np.arange(lp1) + 1.0))[:, None, None]
    c = lambda k: prof[k][:, None]
    or1, or2 = rad["or1"][:, None], rad["or2"][:, None]
    be, dbe, dLv = c("beta"), c("dbeta"), c("dLvisc")
    a = rn * dL * or2 * (R - w0 * c("visc") * (
        D2 + (dLv - be) * D - (dLv * be + 2.0 * dLv * or1 + dL * or2 + dbe + 2.0 * be * or1) * R))
    o1, b = rad["or1"], prof["beta"]
    a[:, 0] = rn * (D[0] - (2.0 * o1[0] + b[0]) * R[0]) if p.ktopv == 1 else rn * R[0]
    a[:, N - 1] = (rn * (D[N - 1] - (2.0 * o1[N - 1] + b[N - 1]) * R[N - 1]) if p.kbotv == 1
                   else rn * R[N - 1])
    if p.n_cheb_max < N:
        a[:, (0, N - 1), p.n_cheb_max:] = 0.0
    a[0] = np.eye(N)   # no l = 0 equation
    return _halve(a, (0, N - 1))


def wp_inverse(p, rad, prof, w0):
    a = _wp_matrix(p, rad, prof, w0)
    rm = np.abs(a).max(axis=2)
    rm[rm == 0] = 1.0
    fr = 1.0 / rm
    a = fr[:, :, None] * a
    cm = np.abs(a).max(axis=1)
    cm[cm == 0] = 1.0
    fc = 1.0 / cm
    a = a * fc[:, None, :]
    out = np.linalg.inv(a) * fc[:, :, None] * fr[:, None, :]
    out[0] = 0.0
    return out


def _wp_matrix(p, rad, prof, w0):
    R, D, D2, D3 = rad["T"], rad["T1"], rad["T2"], rad["T3"]
    N, rn, lp1 = p.n_r_max, rad["rnorm"], p.l_max + 1
    b = slice(1, N - 1)
    dL = (np.arange(lp1) * (np.arange(lp1) + 1.0))[:, None, None]
    c = lambda k: prof[k][b, None]
    o1, o2 = rad["or1"][b, None], rad["or2"][b, None]
    be, dbe, dLv, vi = c("beta"), c("dbeta"), c("dLvisc"), c("visc")
    a = np.zeros((lp1, 2 * N, 2 * N))
    a[:, b, :N] = rn * dL * o2 * (R[b] - w0 * vi * (
        D2[b] + (2.0 * dLv - be / 3.0) * D[b]
        - (dL * o2 + 4.0 / 3.0 * (dLv * be + (3.0 * dLv + be) * o1 + dbe)) * R[b]))
    a[:, b, N:] = rn * w0 * (D[b] - be * R[b])
    a[:, N + 1:2 * N - 1, :N] = rn * dL * o2 * (-D[b] - w0 * vi * (
        -D3[b] + (be - dLv) * D2[b] + (dL * o2 + dbe + dLv * be + 2.0 * (dLv + be) * o1) * D[b]
        - dL * o2 * (2.0 * o1 + dLv + 2.0 / 3.0 * be) * R[b]))
    a[:, N + 1:2 * N - 1, N:] = -rn * w0 * dL * o2 * R[b]
    or1, bt = rad["or1"], prof["beta"]
    a[:, 0, :N] = rn * R[0]
    a[:, N - 1, :N] = rn * R[N - 1]
    a[:, N, :N] = rn * (D2[0] - (2.0 * or1[0] + bt[0]) * D[0]) if p.ktopv == 1 else rn * D[0]
    a[:, 2 * N - 1, :N] = (rn * (D2[N - 1] - (2.0 * or1[N - 1] + bt[N - 1]) * D[N - 1])
                           if p.kbotv == 1 else rn * D[N - 1])
    if p.n_cheb_max < N:
        for row in (0, N - 1, N, 2 * N - 1):
            a[:, row, p.n_cheb_max:N] = 0.0
            a[:, row, N + p.n_cheb_max:] = 0.0
    _halve(a, (0, N - 1, N, 2 * N - 1))
    a[0] = np.eye(2 * N)   # no l = 0 equation
    return a


def affine_operators(p, rad, prof):
    """The S, Z and (w, p) collocation matrices as A0 + w A1 (each is affine in
    the implicit weight w): {"sz": (A0, A1) of shape (2, l_max+1, N, N),
    "wp": (A0, A1) of shape (l_max+1, 2N, 2N)}, for the device rebuild."""
    out = {}
    for key, f in (("sz", lambda w: np.stack([_s_matrix(p, rad, prof, w),
                                               _z_matrix(p, rad, prof, w)])),
                   ("wp", lambda w: _wp_matrix(p, rad, prof, w))):
        a0 = f(0.0)
        out[key] = (a0, f(1.0) - a0)
    return out


def p0_block(p, rad, prof, k4wp, rpad6):
    """Degree-0 block of the (w, p) row operator: p(l=0) and dp/dr(l=0) from the
    pressure half of the right-hand side through the l = 0 pressure inverse."""
    N = p.n_r_max
    opp = np.concatenate([rad["CT"], rad["D1"] @ rad["CT"]])
    mt = np.zeros((k4wp, rpad6))
    mt[N:2 * N, 4 * N:6 * N] = (opp @ p0_inverse(p, rad, prof)).T
    return mt


def integral_row(rad, phys):
    """The Chebyshev integral condition row of a physical profile: its DCT-I
    coefficients (times rnorm, end coefficients halved) through the kernel
    K[j, k] = (1/(1-(k-j)^2) + 1/(1-(k+j)^2)) rnorm/2 for even j + k."""
    n, rn = phys.shape[0], rad["rnorm"]
    work = (rad["C"] @ phys) * rn
    work[0] *= 0.5
    work[n - 1] *= 0.5
    j = np.arange(n, dtype=np.float64)[:, None]
    k = j.T
    d1, d2 = 1.0 - (k - j) ** 2, 1.0 - (k + j) ** 2
    K = np.where(d1 != 0, 1.0 / np.where(d1 != 0, d1, 1.0), 0.0) \
        + np.where(d2 != 0, 1.0 / np.where(d2 != 0, d2, 1.0), 0.0)
    K = np.where((j + k) % 2 == 0, K, 0.0) * 0.5 * rn
    return K @ work


def p0_integral(p, prof) -> bool:
    """True when the l = 0 pressure carries the integral condition."""
    return p.l_anel and prof["ViscHeatFac"] * prof["ThExpNb"] != 0.0


def p0_inverse(p, rad, prof):
    R, D = rad["T"], rad["T1"]
    N, rn, r = p.n_r_max, rad["rnorm"], rad["r"]
    a = np.zeros((N, N))
    a[1:] = rn * (D[1:] - prof["beta"][1:, None] * R[1:])
    if p0_integral(p, prof):
        a[0] = integral_row(rad, prof["ThExpNb"] * prof["ViscHeatFac"] * prof["ogrun"]
                            * prof["alpha0"] * r * r)
    else:
        a[0] = rn * R[0]
    if p.n_cheb_max < N:
        a[0, p.n_cheb_max:] = 0.0
    return np.linalg.inv(_halve(a, (0, N - 1)))


def p0_weights(p, rad, prof):
    """u (N,) with the l = 0 pressure's integral right-hand side u . Re s(l=0)."""
    r = rad["r"]
    return prof["ThExpNb"] * prof["alpha0"] * prof["temp0"] * prof["rho0"] * r * r * rad["q"]


def conduction_state(p, rad, prof):
    """(s0, p0) on the grid: the anelastic conduction state
    (``init_fields.py:190-305``; fixed entropy at both walls)."""
    R, D, D2 = rad["T"], rad["T1"], rad["T2"]
    N, rn, r = p.n_r_max, rad["rnorm"], rad["r"]
    M = np.zeros((2 * N, 2 * N))
    d1 = (prof["beta"] + prof["dLtemp0"] + 2.0 / r + prof["dLkappa"])[:, None]
    M[:N, :N] = rn * p.opr * prof["kappa"][:, None] * (D2 + d1 * D)
    M[N:, :N] = -rn * (prof["rho0"] * p.BuoFac * prof["rgrav"])[:, None] * R
    M[N:, N:] = rn * (D - prof["beta"][:, None] * R)
    M[0] = 0.0
    M[0, :N] = rn * R[0]
    M[N - 1] = 0.0
    M[N - 1, :N] = rn * R[N - 1]
    if prof["ViscHeatFac"] * prof["ThExpNb"] != 0.0:
        M[N, N:] = integral_row(rad, prof["ThExpNb"] * prof["ViscHeatFac"] * prof["ogrun"]
                                * prof["alpha0"] * r * r)
        M[N, :N] = integral_row(rad, -prof["ThExpNb"] * prof["alpha0"] * prof["temp0"]
                                * prof["rho0"] * r * r)
    else:
        M[N] = 0.0
        M[N, N:] = rn * R[0]
    _halve(M, (0, N - 1, N, 2 * N - 1))
    rhs = np.zeros(2 * N)
    rhs[N - 1] = math.sqrt(4.0 * math.pi)
    fac = 1.0 / np.abs(M).max(axis=1)
    x = np.linalg.solve(fac[:, None] * M, fac * rhs)
    return rad["C"] @ x[:N], rad["C"] @ x[N:]


def am_setup(p, rad, prof):
    """Angular-momentum correction profiles and factors (``update_z.py:57-66``,
    ``pre_calculations.py:60-105``): rho0 r^2 and its first two derivatives,
    r^2 times the radial weights, and the factors of the axial and equatorial
    corrections (8/3 pi y10 / (c_moi_oc y10), 8/3 pi / c_moi_oc)."""
    r, q = rad["r"], rad["q"]
    rho0, be, dbe = prof["rho0"], prof["beta"], prof["dbeta"]
    r2 = r * r
    c_moi = 8.0 / 3.0 * math.pi * float((r2 * r2 * rho0) @ q)
    y10 = 0.5 * math.sqrt(3.0 / math.pi)
    return dict(r2q=r2 * q, pz=rho0 * r2, pdz=rho0 * (2.0 * r + r2 * be),
                pd2z=rho0 * (2.0 + 4.0 * be * r + dbe * r2 + be * be * r2),
                cz=8.0 / 3.0 * math.pi * y10 / (c_moi * y10), ce=8.0 / 3.0 * math.pi / c_moi)
