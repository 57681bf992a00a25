"""CPU FP64 restatement of the reference's finite-difference radial scheme and
its banded implicit operators.  TEST INFRASTRUCTURE: only tests/, smoke() and
bench.py's CPU baseline may use this module.

Follows, in order:
* ``finite_differences.py:20-59``   Fornberg weights (``populate_fd_weights``)
* ``finite_differences.py:62-127``  stretched FD grid (``_get_fd_grid``)
* ``finite_differences.py:130-276`` stencil coefficients (``_get_fd_coeffs``)
* ``finite_differences.py:279-331`` derivative matrices (``_build_fd_der_mats``)
* ``radial_functions.py:66-160``    polytropic anelastic background
* ``update_s.py:66-190``            entropy operator, FD branch (band kl=ku=hw)
* ``update_z.py:89-240``            toroidal operator, FD branch
* ``update_w_doublecurl.py:67-307`` double-curl poloidal operator, FD branch

Pins (tests/test_oracle_fd.py): Fornberg weights reproduce derivatives of
polynomials exactly up to the stencil order; the derivative matrices are exact
on polynomials of the order of their stencils; the stretched grid has the
reference's end points, monotonicity and spacing ratio; the polytropic
profiles satisfy rho0 = temp0^polind, beta = d ln rho0/dr (checked by finite
differences of the closed form) and the density contrast exp(strat) between
the boundaries; every operator fits its band (dense_to_band_storage accepts
it) and its LINPACK factorisation reproduces scipy's dgbtrf pivots.
"""

from __future__ import annotations

import math

import numpy as np


def fornberg(z: float, x: np.ndarray, nd: int, m: int) -> np.ndarray:
    """Weights c[j, k] of derivative k at z from values at x[0..nd]
    (``finite_differences.py:20-59``, Fornberg 1988)."""
    c = np.zeros((nd + 1, m + 1))
    c1 = 1.0
    c4 = x[0] - z
    c[0, 0] = 1.0
    for i in range(1, nd + 1):
        mn = min(i, m)
        c2 = 1.0
        c5 = c4
        c4 = x[i] - z
        for j in range(i):
            c3 = x[i] - x[j]
            c2 = c2 * c3
            if j == i - 1:
                for k in range(mn, 0, -1):
                    c[i, k] = c1 * (k * c[i - 1, k - 1] - c5 * c[i - 1, k]) / c2
                c[i, 0] = -c1 * c5 * c[i - 1, 0] / c2
            for k in range(mn, 0, -1):
                c[j, k] = (c4 * c[j, k] - k * c[j, k - 1]) / c3
            c[j, 0] = c4 * c[j, 0] / c3
        c1 = c2
    return c


def fd_grid(n: int, ricb: float, rcmb: float, ratio1: float, ratio2: float) -> np.ndarray:
    """Radial FD grid r[0] = rcmb .. r[n-1] = ricb (``finite_differences.py:62-127``)."""
    r = np.zeros(n)
    r[0] = rcmb
    if ratio2 == 1.0:
        dr = 1.0 / (n - 1)
        for i in range(1, n):
            r[i] = r[i - 1] - dr
    else:
        nb = int((n - 1) / (2.0 * (1.0 + ratio1)))
        full_sphere = abs(ricb) <= 10.0 * np.finfo(np.float64).eps
        dr_after = math.exp(math.log(ratio2) / nb)
        dr_before = 1.0
        for _ in range(nb):
            dr_before *= dr_after
        if full_sphere:
            n_bulk = n - 1 - nb
            dr_before = 1.0 / (n_bulk + dr_after * (1.0 - dr_before) / (1.0 - dr_after))
        else:
            n_bulk = n - 1 - 2 * nb
            dr_before = 1.0 / (n_bulk + 2.0 * dr_after * (1.0 - dr_before) / (1.0 - dr_after))
        for _ in range(nb):
            dr_before *= dr_after
        for i in range(1, nb + 1):
            r[i] = r[i - 1] - dr_before
            dr_before /= dr_after
        for i in range(n_bulk):
            r[nb + 1 + i] = r[nb + i] - dr_before
        if not full_sphere:
            for i in range(nb):
                dr_before *= dr_after
                r[nb + n_bulk + 1 + i] = r[nb + n_bulk + i] - dr_before
    if r[n - 1] != ricb:
        r[n - 1] = ricb
    return r


def fd_coeffs(r: np.ndarray, order: int, order_boundary: int) -> dict:
    """Bulk and boundary stencils (``finite_differences.py:130-276``)."""
    n = r.shape[0]
    h = order // 2
    dr = np.zeros((n, order + 1))
    ddr = np.zeros((n, order + 1))
    if order == 2:
        for i in range(n):
            if i in (0, n - 1):
                d = (r[1] - r[0]) if i == 0 else (r[n - 1] - r[n - 2])
                dr[i] = (-0.5 / d, 0.0, 0.5 / d)
                ddr[i] = (1.0 / d / d, -2.0 / d / d, 1.0 / d / d)
            else:
                c = fornberg(0.0, np.array([r[i - 1] - r[i], 0.0, r[i + 1] - r[i]]), 2, 2)
                dr[i], ddr[i] = c[:, 1], c[:, 2]
    else:
        for i in range(h, n - h):
            c = fornberg(0.0, np.array([r[i - h + o] - r[i] for o in range(order + 1)]), order, order)
            dr[i], ddr[i] = c[:, 1], c[:, 2]
    ob = order_boundary
    dr_top = np.zeros((h, ob + 1)); dr_bot = np.zeros((h, ob + 1))
    ddr_top = np.zeros((h, ob + 2)); ddr_bot = np.zeros((h, ob + 2))
    for i in range(h):
        dr_top[i] = fornberg(0.0, np.array([r[o] - r[i] for o in range(ob + 1)]), ob, ob)[:, 1]
        dr_bot[i] = fornberg(0.0, np.array([r[n - 1 - o] - r[n - 1 - i] for o in range(ob + 1)]),
                             ob, ob)[:, 1]
        ddr_top[i] = fornberg(0.0, np.array([r[o] - r[i] for o in range(ob + 2)]), ob + 1, ob + 1)[:, 2]
        ddr_bot[i] = fornberg(0.0, np.array([r[n - 1 - o] - r[n - 1 - i] for o in range(ob + 2)]),
                              ob + 1, ob + 1)[:, 2]
    dddr = np.zeros((n, order + 3))
    ddddr = np.zeros((n, order + 3))
    if order == 2:
        for i in range(1, n - 1):
            sp = np.zeros(5)
            for o in range(5):
                if i == 1 and o == 0:
                    sp[o] = 2.0 * (r[0] - r[1])
                elif i == n - 2 and o == 4:
                    sp[o] = 2.0 * (r[n - 1] - r[n - 2])
                else:
                    sp[o] = r[i - 2 + o] - r[i]
            c = fornberg(0.0, sp, 4, 4)
            dddr[i], ddddr[i] = c[:, 3], c[:, 4]
    else:
        for i in range(1 + h, n - h - 1):
            c = fornberg(0.0, np.array([r[i - h - 1 + o] - r[i] for o in range(order + 3)]),
                         order + 2, order + 2)
            dddr[i], ddddr[i] = c[:, 3], c[:, 4]
    d3t = np.zeros((h + 1, ob + 3)); d3b = np.zeros((h + 1, ob + 3))
    d4t = np.zeros((h + 1, ob + 4)); d4b = np.zeros((h + 1, ob + 4))
    for i in range(h + 1):
        d3t[i] = fornberg(0.0, np.array([r[o] - r[i] for o in range(ob + 3)]), ob + 2, ob + 2)[:, 3]
        d3b[i] = fornberg(0.0, np.array([r[n - 1 - o] - r[n - 1 - i] for o in range(ob + 3)]),
                          ob + 2, ob + 2)[:, 3]
        d4t[i] = fornberg(0.0, np.array([r[o] - r[i] for o in range(ob + 4)]), ob + 3, 4)[:, 4]
        d4b[i] = fornberg(0.0, np.array([r[n - 1 - o] - r[n - 1 - i] for o in range(ob + 4)]),
                          ob + 3, 4)[:, 4]
    return dict(dr=dr, ddr=ddr, dddr=dddr, ddddr=ddddr, dr_top=dr_top, dr_bot=dr_bot,
                ddr_top=ddr_top, ddr_bot=ddr_bot, dddr_top=d3t, dddr_bot=d3b,
                ddddr_top=d4t, ddddr_bot=d4b)


def fd_der_mats(r: np.ndarray, c: dict, order: int, order_boundary: int):
    """(rMat, drMat, d2rMat, d3rMat, d4rMat) dense (``finite_differences.py:279-331``)."""
    n = r.shape[0]
    h = order // 2
    ob = order_boundary
    R = np.eye(n)
    D1, D2, D3, D4 = (np.zeros((n, n)) for _ in range(4))
    for i in range(h, n - h):
        D1[i, i - h:i + h + 1] = c["dr"][i]
        D2[i, i - h:i + h + 1] = c["ddr"][i]
    for i in range(1 + h, n - h - 1):
        D3[i, i - h - 1:i + h + 2] = c["dddr"][i]
        D4[i, i - h - 1:i + h + 2] = c["ddddr"][i]
    for i in range(h):
        D1[i, :ob + 1] = c["dr_top"][i]
        D1[n - 1 - i, n - ob - 1:n] = c["dr_bot"][i][::-1]
        D2[i, :ob + 2] = c["ddr_top"][i]
        D2[n - 1 - i, n - ob - 2:n] = c["ddr_bot"][i][::-1]
    for i in range(h + 1):
        D3[i, :ob + 3] = c["dddr_top"][i]
        D3[n - 1 - i, n - ob - 3:n] = c["dddr_bot"][i][::-1]
        D4[i, :ob + 4] = c["ddddr_top"][i]
        D4[n - 1 - i, n - ob - 4:n] = c["ddddr_bot"][i][::-1]
    return R, D1, D2, D3, D4


def polytropic(r: np.ndarray, rcmb: float, ricb: float, radratio: float, strat: float,
               polind: float, g0: float, g1: float, g2: float) -> dict:
    """Polytropic reference state (``radial_functions.py:66-160``)."""
    or1 = 1.0 / r
    or2, or3 = or1 * or1, or1 ** 3
    or4 = or2 * or2
    rgrav = g0 + g1 * r / rcmb + g2 * rcmb ** 2 * or2
    DissNb = ((math.exp(strat / polind) - 1.0) / (rcmb - ricb)
              / (g0 + 0.5 * g1 * (1.0 + radratio) + g2 / radratio))
    temp0 = (-DissNb * (g0 * r + 0.5 * g1 * r ** 2 / rcmb - g2 * rcmb ** 2 * or1)
             + 1.0 + DissNb * rcmb * (g0 + 0.5 * g1 - g2))
    rho0 = temp0 ** polind
    beta = -polind * DissNb * rgrav / temp0
    dgrav = g1 / rcmb - 2.0 * g2 * rcmb ** 2 * or3
    dbeta = -polind * DissNb / temp0 ** 2 * (dgrav * temp0 + DissNb * rgrav ** 2)
    ddbeta = -polind * DissNb / temp0 ** 3 * (
        temp0 * (6.0 * g2 * rcmb ** 2 * or4 * temp0 + DissNb * rgrav * dgrav)
        + 2.0 * DissNb * rgrav * (DissNb * rgrav ** 2 + dgrav * temp0))
    dtemp0 = -DissNb * rgrav
    dLtemp0 = dtemp0 / temp0
    n = r.shape[0]
    return dict(or1=or1, or2=or2, rgrav=rgrav, temp0=temp0, rho0=rho0, beta=beta,
                dbeta=dbeta, ddbeta=ddbeta, dLtemp0=dLtemp0, orho1=1.0 / rho0,
                visc=np.ones(n), dLvisc=np.zeros(n), ddLvisc=np.zeros(n),
                kappa=np.ones(n), dLkappa=np.zeros(n), DissNb=DissNb)


def _precond_rows(dat):
    fac = 1.0 / np.abs(dat).max(axis=1)
    return fac[:, None] * dat, fac


def band_widths(order: int, order_boundary: int):
    """(kl for S and Z with no-slip walls, kl for W): ``update_s.py:103-108``,
    ``update_z.py:121-127``, ``update_w_doublecurl.py:112-118`` (ku = kl)."""
    if order <= 2 and order_boundary <= 2:
        hs = 1
        nbw = order + 3
    else:
        hs = max(order // 2, order_boundary)
        nbw = max(order + 3, 2 * order_boundary + 3)
    return hs, (nbw - 1) // 2


def s_operator(l: int, wimp: float, opr: float, mats, prof):
    """Row-preconditioned entropy operator of degree l (``update_s.py:146-163``)."""
    R, D1, D2 = mats[0], mats[1], mats[2]
    n = R.shape[0]
    dL = float(l * (l + 1))
    d1c = (prof["beta"] + prof["dLtemp0"] + 2.0 * prof["or1"] + prof["dLkappa"])[:, None]
    dat = R - wimp * opr * 1.0 * prof["kappa"][:, None] * (D2 + d1c * D1 - dL * prof["or2"][:, None] * R)
    dat[0, :] = R[0, :]
    dat[n - 1, :] = R[n - 1, :]
    return _precond_rows(dat)


def z_operator(l: int, wimp: float, mats, prof, ktopv=2, kbotv=2):
    """Row-preconditioned toroidal operator of degree l >= 1 (``update_z.py:160-190``)."""
    R, D1, D2 = mats[0], mats[1], mats[2]
    n = R.shape[0]
    dL = float(l * (l + 1))
    beta, dbeta, dLvisc = prof["beta"][:, None], prof["dbeta"][:, None], prof["dLvisc"][:, None]
    or1, or2 = prof["or1"][:, None], prof["or2"][:, None]
    dat = dL * or2 * (R - wimp * 1.0 * prof["visc"][:, None] * (
        D2 + (dLvisc - beta) * D1
        - (dLvisc * beta + 2.0 * dLvisc * or1 + dL * or2 + dbeta + 2.0 * beta * or1) * R))
    o1, b = prof["or1"], prof["beta"]
    dat[0, :] = (D1[0, :] - (2.0 * o1[0] + b[0]) * R[0, :]) if ktopv == 1 else R[0, :]
    dat[n - 1, :] = (D1[n - 1, :] - (2.0 * o1[n - 1] + b[n - 1]) * R[n - 1, :]) if kbotv == 1 \
        else R[n - 1, :]
    return _precond_rows(dat)


def w_operator(l: int, wimp: float, mats, prof, ktopv=2, kbotv=2):
    """Row- and column-preconditioned double-curl operator of degree l >= 1
    (``update_w_doublecurl.py:153-242``); returns (dat, fac_row, fac_col)."""
    R, D1, D2, D3, D4 = mats
    n = R.shape[0]
    dL = float(l * (l + 1))
    b = slice(2, n - 2)
    o1, o2 = prof["or1"][b], prof["or2"][b]
    be, dbe, ddbe = prof["beta"][b], prof["dbeta"][b], prof["ddbeta"][b]
    vi, dv, ddv, orho = prof["visc"][b], prof["dLvisc"][b], prof["ddLvisc"][b], prof["orho1"][b]
    c3 = 2.0 * (dv - be)
    c2 = (ddv - 2.0 * dbe + dv ** 2 + be ** 2 - 3.0 * dv * be - 2.0 * o1 * (dv + be)) - 2.0 * dL * o2
    c1 = (-ddbe - dbe * (2.0 * dv - be + 2.0 * o1) - ddv * (be + 2.0 * o1)
          + be ** 2 * (dv + 2.0 * o1) - be * (dv ** 2 - 2.0 * o2)
          - 2.0 * dv * o1 * (dv - o1)) + dL * 2.0 * (2.0 * o1 + be - dv) * o2
    c0 = dL * o2 * ((2.0 * dbe + ddv + dv ** 2 - (2.0 / 3.0) * be ** 2 + dv * be
                     + 2.0 * o1 * (2.0 * dv - be - 3.0 * o1)) + dL * o2)
    dat = np.zeros((n, n))
    old = (-dL * o2 * orho)[:, None] * (D2[b] - be[:, None] * D1[b] - dL * o2[:, None] * R[b])
    imp = (orho * 1.0 * vi * dL * o2)[:, None] * (
        D4[b] + c3[:, None] * D3[b] + c2[:, None] * D2[b] + c1[:, None] * D1[b] + c0[:, None] * R[b])
    dat[b] = old + wimp * imp
    dat[0, :] = R[0, :]
    dat[n - 1, :] = R[n - 1, :]
    ob, bb = prof["or1"], prof["beta"]
    dat[1, :] = (D2[0, :] - (2.0 * ob[0] + bb[0]) * D1[0, :]) if ktopv == 1 else D1[0, :]
    dat[n - 2, :] = (D2[n - 1, :] - (2.0 * ob[n - 1] + bb[n - 1]) * D1[n - 1, :]) if kbotv == 1 \
        else D1[n - 1, :]
    fr = 1.0 / np.abs(dat).max(axis=1)
    dat = fr[:, None] * dat
    fc = 1.0 / np.abs(dat).max(axis=0)
    return dat * fc[None, :], fr, fc
