"""Per-l implicit operator matrices for the Boussinesq Chebyshev step (oracle).

Restates ``update_s.py:66-188`` (sMat), ``update_z.py:89-240`` (zMat),
``update_wp.py:80-170`` (p0Mat), ``update_wp.py:197-368`` (coupled wpMat with
row and column preconditioning) and ``update_b.py:98-298`` (insulating
bMat/jMat), each with Boussinesq profiles (visc=kappa=lambda=1,
beta=dLvisc=dbeta=0, hdif=1).  As in the reference, the stored operator is
the preconditioned inverse re-scaled so that ``inv_by_l[l] @ rhs`` gives the
Chebyshev coefficients, and the l=0 inverse is zero for Z, WP, B and J.

Pinned by ``tests/test_oracle_matrices.py``: every per-l inverse reproduces
the identity against an independently written matrix (the reference's own
``tests/test_implicit_matrices.py:22-365`` pattern), the p0/ps_cond pair
reproduces the analytic Boussinesq conduction state s0 = a + b/r.
"""

from __future__ import annotations

import numpy as np

from .params import Params


def _precond_rows(dat):
    fac = 1.0 / np.abs(dat).max(axis=1)
    return fac[:, None] * dat, fac


def _bfac_cols(dat, cols, f=0.5):
    """Boundary factor on the first/last outer columns: 1/2 for Chebyshev
    coefficients, 1 on the finite-difference grid (boundary_fac)."""
    for c in cols:
        dat[:, c] *= f
    return dat


def build_s(p: Params, rad, mats, wimp_lin0):
    """sMat per l (``update_s.py:66-188``), returns inv_by_l (l_max+1, N, N)."""
    R, D, D2, _ = mats
    N = p.n_r_max
    rn = rad["rnorm"]
    or1 = rad["or1"][:, None]; or2 = rad["or2"][:, None]
    out = np.zeros((p.l_max + 1, N, N))
    for l in range(p.l_max + 1):
        dL = float(l * (l + 1))
        dat = rn * (R - wimp_lin0 * p.opr * (D2 + 2.0 * or1 * D - dL * or2 * R))
        dat[0] = rn * R[0]
        dat[N - 1] = rn * R[N - 1]
        if p.n_cheb_max < N:
            dat[0, p.n_cheb_max:] = 0.0
            dat[N - 1, p.n_cheb_max:] = 0.0
        _bfac_cols(dat, (0, N - 1))
        datp, fac = _precond_rows(dat)
        out[l] = np.linalg.inv(datp) * fac[None, :]
    return out


def build_z(p: Params, rad, mats, wimp_lin0):
    """zMat per l>=1 (``update_z.py:89-240``), no-slip or stress-free rows."""
    R, D, D2, _ = mats
    N = p.n_r_max
    rn = rad["rnorm"]
    or1 = rad["or1"]; or2c = rad["or2"][:, None]
    out = np.zeros((p.l_max + 1, N, N))
    for l in range(1, p.l_max + 1):
        dL = float(l * (l + 1))
        dat = rn * dL * or2c * (R - wimp_lin0 * (D2 - dL * or2c * R))
        dat[0] = rn * (D[0] - 2.0 * or1[0] * R[0]) if p.ktopv == 1 else rn * R[0]
        dat[N - 1] = (rn * (D[N - 1] - 2.0 * or1[N - 1] * R[N - 1]) if p.kbotv == 1
                      else rn * R[N - 1])
        if p.n_cheb_max < N:
            dat[0, p.n_cheb_max:] = 0.0
            dat[N - 1, p.n_cheb_max:] = 0.0
        _bfac_cols(dat, (0, N - 1))
        datp, fac = _precond_rows(dat)
        out[l] = np.linalg.inv(datp) * fac[None, :]
    return out


def build_z10(p: Params, rad, mats, wimp_lin0, c_dt_z10_ic):
    """z10Mat of the rotating inner core (``update_z.py:244-298``, Boussinesq,
    time-integrated rotation): the l = 1 zMat bulk rows, z = 0 at the CMB, and
    at the ICB the angular-momentum balance c_dt_z10_ic z + wimp_lin0 (2 z / r -
    dz/dr); returns the (N, N) inverse with the row scaling folded in."""
    R, D, D2, _ = mats
    N = p.n_r_max
    rn = rad["rnorm"]
    or1 = rad["or1"]; or2c = rad["or2"][:, None]
    dL = 2.0
    dat = rn * dL * or2c * (R - wimp_lin0 * (D2 - dL * or2c * R))
    dat[0] = rn * R[0]
    dat[N - 1] = rn * (c_dt_z10_ic * R[N - 1]
                       + wimp_lin0 * (2.0 * or1[N - 1] * R[N - 1] - D[N - 1]))
    if p.n_cheb_max < N:
        dat[0, p.n_cheb_max:] = 0.0
        dat[N - 1, p.n_cheb_max:] = 0.0
    _bfac_cols(dat, (0, N - 1), rad.get("bfac", 0.5))
    datp, fac = _precond_rows(dat)
    return np.linalg.inv(datp) * fac[None, :]


def build_p0(p: Params, rad, mats):
    """l=0 pressure matrix inverse (``update_wp.py:80-170``, Chebyshev, Boussinesq)."""
    R, D, _, _ = mats
    N = p.n_r_max
    rn = rad["rnorm"]
    dat = np.zeros((N, N))
    dat[1:] = rn * D[1:]
    dat[0] = rn * R[0]
    if p.n_cheb_max < N:
        dat[0, p.n_cheb_max:] = 0.0
    _bfac_cols(dat, (0, N - 1))
    return np.linalg.inv(dat)


def build_wp(p: Params, rad, mats, wimp_lin0):
    """Coupled 2N×2N (w, p) matrices (``update_wp.py:197-368``, Boussinesq)."""
    R, D, D2, D3 = mats
    N = p.n_r_max
    rn = rad["rnorm"]
    or1 = rad["or1"]; or2 = rad["or2"]
    o2 = or2[1:N - 1, None]
    o1 = or1[1:N - 1, None]
    out = np.zeros((p.l_max + 1, 2 * N, 2 * N))
    for l in range(1, p.l_max + 1):
        dL = float(l * (l + 1))
        dat = np.zeros((2 * N, 2 * N))
        dat[1:N - 1, :N] = rn * dL * o2 * (R[1:N - 1] - wimp_lin0 * (D2[1:N - 1] - dL * o2 * R[1:N - 1]))
        dat[1:N - 1, N:] = rn * wimp_lin0 * D[1:N - 1]
        dat[N + 1:2 * N - 1, :N] = rn * dL * o2 * (
            -D[1:N - 1] - wimp_lin0 * (-D3[1:N - 1] + dL * o2 * D[1:N - 1]
                                       - dL * o2 * (2.0 * o1) * R[1:N - 1]))
        dat[N + 1:2 * N - 1, N:] = -rn * wimp_lin0 * dL * o2 * R[1:N - 1]
        dat[0, :N] = rn * R[0]
        dat[N - 1, :N] = rn * R[N - 1]
        dat[N, :N] = rn * (D2[0] - 2.0 * or1[0] * D[0]) if p.ktopv == 1 else rn * D[0]
        dat[2 * N - 1, :N] = (rn * (D2[N - 1] - 2.0 * or1[N - 1] * D[N - 1]) if p.kbotv == 1
                              else rn * D[N - 1])
        if p.n_cheb_max < N:
            for row in (0, N - 1, N, 2 * N - 1):
                dat[row, p.n_cheb_max:N] = 0.0
                dat[row, N + p.n_cheb_max:] = 0.0
        _bfac_cols(dat, (0, N - 1, N, 2 * N - 1))
        row_max = np.abs(dat).max(axis=1)
        row_max[row_max == 0] = 1.0
        fr = 1.0 / row_max
        dat = fr[:, None] * dat
        col_max = np.abs(dat).max(axis=0)
        col_max[col_max == 0] = 1.0
        fc = 1.0 / col_max
        dat = dat * fc[None, :]
        out[l] = np.linalg.inv(dat) * fc[:, None] * fr[None, :]
    return out


def build_bj(p: Params, rad, mats, wimp_lin0):
    """Insulating bMat, jMat (``update_b.py:98-298``); returns (b_inv, j_inv)."""
    R, D, D2, _ = mats
    N = p.n_r_max
    rn = rad["rnorm"]
    or1 = rad["or1"]; or2 = rad["or2"]
    o2 = or2[1:N - 1, None]
    b_out = np.zeros((p.l_max + 1, N, N))
    j_out = np.zeros((p.l_max + 1, N, N))
    for l in range(1, p.l_max + 1):
        dL = float(l * (l + 1))
        for which in ("b", "j"):
            dat = np.zeros((N, N))
            dat[1:N - 1] = rn * dL * o2 * (R[1:N - 1] - wimp_lin0 * p.opm * (D2[1:N - 1] - dL * o2 * R[1:N - 1]))
            if which == "b":
                dat[0] = rn * (D[0] + float(l) * or1[0] * R[0])
                dat[N - 1] = rn * (D[N - 1] - float(l + 1) * or1[N - 1] * R[N - 1])
            else:
                dat[0] = rn * R[0]
                dat[N - 1] = rn * R[N - 1]
            if p.n_cheb_max < N:
                dat[0, p.n_cheb_max:] = 0.0
                dat[N - 1, p.n_cheb_max:] = 0.0
            _bfac_cols(dat, (0, N - 1))
            datp, fac = _precond_rows(dat)
            inv = np.linalg.inv(datp) * fac[None, :]
            (b_out if which == "b" else j_out)[l] = inv
    return b_out, j_out


def build_bj_coupled(p: Params, rad, mats, ic, wimp_lin0):
    """Coupled outer-core + inner-core bMat, jMat of a conducting inner core
    (``update_b.py:301-513``, Chebyshev): rows 0 CMB condition, 1..N-2 outer
    diffusion, N-1 continuity at the ICB, N derivative match at the ICB
    (sigma_ratio on the outer dj/dr for j), N+1..NT-2 inner diffusion, NT-1 the
    centre limit; columns 0..N-1 outer Chebyshev, N..NT-1 inner even
    Chebyshev coefficients.  Returns (b_inv, j_inv), (l_max+1, NT, NT)."""
    R, D, D2, _ = mats
    N, Nic, NT = p.n_r_max, p.n_r_ic_max, p.n_r_tot
    nci = p.n_cheb_ic_max
    rn, cn = rad["rnorm"], ic["cnorm"]
    or1, or2 = rad["or1"], rad["or2"]
    o2 = or2[1:N - 1, None]
    or2_icb = or2[N - 1]
    cheb, dcheb, d2cheb = ic["cheb"], ic["dcheb"], ic["d2cheb"]
    bulk = slice(1, Nic - 1)
    ic_eval, dic_eval, d2ic_eval = cheb[:, bulk].T, dcheb[:, bulk].T, d2cheb[:, bulk].T
    or_ic = ic["O_r_ic"][bulk, None]
    b_out = np.zeros((p.l_max + 1, NT, NT))
    j_out = np.zeros((p.l_max + 1, NT, NT))
    for l in range(1, p.l_max + 1):
        dL = float(l * (l + 1))
        for which in ("b", "j"):
            dat = np.zeros((NT, NT))
            dat[1:N - 1, :N] = rn * dL * o2 * (R[1:N - 1] - wimp_lin0 * p.opm
                                              * (D2[1:N - 1] - dL * o2 * R[1:N - 1]))
            dat[0, :N] = rn * (D[0] + float(l) * or1[0] * R[0]) if which == "b" else rn * R[0]
            dat[N - 1, :N] = rn * R[N - 1]
            dat[N - 1, N:N + nci] = -cn * cheb[:nci, 0]
            dat[N, :N] = rn * (1.0 if which == "b" else p.sigma_ratio) * D[N - 1]
            dat[N, N:N + nci] = -cn * (dcheb[:nci, 0] + float(l + 1) * or1[N - 1] * cheb[:nci, 0])
            if p.n_cheb_max < N:
                dat[(0, N - 1, N), p.n_cheb_max:N] = 0.0
            dat[N + 1:NT - 1, N:] = cn * dL * or2_icb * (
                ic_eval - wimp_lin0 * p.opm * p.O_sr * (d2ic_eval + 2.0 * float(l + 1) * or_ic * dic_eval))
            dat[NT - 1, N:] = cn * dL * or2_icb * (
                cheb[:, Nic - 1] - wimp_lin0 * p.opm * p.O_sr * (1.0 + 2.0 * float(l + 1))
                * d2cheb[:, Nic - 1])
            _bfac_cols(dat, (0, N - 1), rad.get("bfac", 0.5))
            dat[:, N] *= 0.5
            dat[:, NT - 1] *= 0.5
            datp, fac = _precond_rows(dat)
            (b_out if which == "b" else j_out)[l] = np.linalg.inv(datp) * fac[None, :]
    return b_out, j_out


def ps_cond(p: Params, rad, mats):
    """Boussinesq conduction state (s0, p0) on the grid (``init_fields.py:108-180``)."""
    from .algebra import prepare_mat, solve_mat
    from .radial import costf
    R, D, D2, _ = mats
    N = p.n_r_max
    rn = rad["rnorm"]
    or1 = rad["or1"]
    M = np.zeros((2 * N, 2 * N))
    M[:N, :N] = rn * p.opr * (D2 + (2.0 * or1)[:, None] * D)
    M[N:, :N] = -rn * (p.BuoFac * rad["rgrav"])[:, None] * R
    M[N:, N:] = rn * D
    M[0, :N] = rn * R[0]; M[0, N:] = 0.0
    M[N - 1, :N] = rn * R[N - 1]; M[N - 1, N:] = 0.0
    M[N, :N] = 0.0; M[N, N:] = rn * R[0]
    for c in (0, N - 1, N, 2 * N - 1):
        M[:, c] *= 0.5
    fac = 1.0 / np.abs(M).max(axis=1)
    M = M * fac[:, None]
    a, ip, info = prepare_mat(M)
    assert info == 0, f"Singular ps0Mat, info={info}"
    rhs = np.zeros(2 * N)
    rhs[N - 1] = np.sqrt(4.0 * np.pi)
    x = solve_mat(a, ip, fac * rhs)
    return costf(x[:N]), costf(x[N:])
