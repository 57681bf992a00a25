"""Conducting inner core: the coupled outer-core + inner-core B and J operators.

With ``sigma_ratio > 0`` the poloidal and toroidal field equations of the
outer core and the inner core are one linear system per degree l over the
n_r_tot = n_r_max + n_r_ic_max coefficients (outer Chebyshev, then inner even
Chebyshev; reference ``update_b.py:301-513``, Chebyshev branch): rows 0 the
CMB condition, 1..N-2 outer diffusion, N-1 continuity at the ICB, N the
derivative match at the ICB (sigma_ratio on the outer dj/dr for J), N+1..NT-2
inner diffusion, NT-1 the limit at the centre; boundary columns halved, rows
scaled by their largest entry.

The matrices (and the z(l=1, m=0) matrix of a rotating core) are affine in
the implicit weight: ``coupled_affine`` / ``z10_affine`` assemble A0 and A1
once on the host, and every rebuild forms A0 + w A1 on the device, inverts all
degrees in one launch and folds the DCT-I and the radial derivatives of both
shells into the two row-operator matrices of each system: outer outputs
(f, f', f'') on the n_r_max levels and inner outputs on the n_r_ic_max levels,
both acting on the coupled right-hand side (reference ``update_b.py:634-748``).
``coupled_inverses`` / ``row_operators`` / ``z10_inverse`` are the host
equivalents (checks).
"""
from __future__ import annotations

import math

import numpy as np


def coupled_inverses(p, rad: dict, ic: dict, wimp_lin0: float):
    """(b_inv, j_inv): (l_max + 1, NT, NT) inverses with the row scaling folded
    in (inverse times diag(row factors)); degree 0 is zero."""
    out = []
    for a in coupled_matrices(p, rad, ic, wimp_lin0):
        fac = 1.0 / np.abs(a).max(axis=2)
        inv = np.linalg.inv(fac[:, :, None] * a) * fac[:, None, :]
        inv[0] = 0.0
        out.append(inv)
    return tuple(out)


def coupled_affine(p, rad: dict, ic: dict):
    """The coupled B and J matrices as A0 + w A1 (affine in the implicit weight):
    ((A0_b, A1_b), (A0_j, A1_j)), each (l_max + 1, NT, NT), degree 0 the identity."""
    m0, m1 = coupled_matrices(p, rad, ic, 0.0), coupled_matrices(p, rad, ic, 1.0)
    return tuple((a, b - a) for a, b in zip(m0, m1))


def coupled_matrices(p, rad: dict, ic: dict, wimp_lin0: float):
    """(b, j): (l_max + 1, NT, NT) collocation matrices of the coupled outer and
    inner systems (degree 0: the identity, no equation)."""
    R, D, D2 = rad["T"], rad["T1"], rad["T2"]
    N, Nic, NT = p.n_r_max, p.n_r_ic_max, p.n_r_tot
    nci = p.n_cheb_ic_max
    rn, cn = rad["rnorm"], ic["cnorm"]
    or1, or2 = rad["or1"], rad["or2"]
    o2 = or2[1:N - 1, None]
    or2_icb = or2[N - 1]
    cheb, dcheb, d2cheb = ic["cheb"], ic["dcheb"], ic["d2cheb"]
    inner = slice(1, Nic - 1)
    c_in, dc_in, d2c_in = cheb[:, inner].T, dcheb[:, inner].T, d2cheb[:, inner].T
    orc = ic["O_r_ic"][inner, None]
    out = {k: np.zeros((p.l_max + 1, NT, NT)) for k in ("b", "j")}
    for k in ("b", "j"):
        out[k][0] = np.eye(NT)
    for l in range(1, p.l_max + 1):
        dL = float(l * (l + 1))
        for key in ("b", "j"):
            a = np.zeros((NT, NT))
            a[1:N - 1, :N] = rn * dL * o2 * (R[1:N - 1] - wimp_lin0 * p.opm
                                            * (D2[1:N - 1] - dL * o2 * R[1:N - 1]))
            a[0, :N] = rn * (D[0] + l * or1[0] * R[0]) if key == "b" else rn * R[0]
            a[N - 1, :N] = rn * R[N - 1]
            a[N - 1, N:N + nci] = -cn * cheb[:nci, 0]
            a[N, :N] = rn * (1.0 if key == "b" else p.sigma_ratio) * D[N - 1]
            a[N, N:N + nci] = -cn * (dcheb[:nci, 0] + (l + 1) * or1[N - 1] * cheb[:nci, 0])
            if p.n_cheb_max < N:
                a[(0, N - 1, N), p.n_cheb_max:N] = 0.0
            a[N + 1:NT - 1, N:] = cn * dL * or2_icb * (
                c_in - wimp_lin0 * p.opm * p.O_sr * (d2c_in + 2.0 * (l + 1) * orc * dc_in))
            a[NT - 1, N:] = cn * dL * or2_icb * (
                cheb[:, Nic - 1] - wimp_lin0 * p.opm * p.O_sr * (1.0 + 2.0 * (l + 1))
                * d2cheb[:, Nic - 1])
            for c in (0, N - 1):      # outer boundary factor (1 on the FD grid)
                a[:, c] *= rad.get("bfac", 0.5)
            for c in (N, NT - 1):     # inner even-Chebyshev end coefficients
                a[:, c] *= 0.5
            out[key][l] = a
    return out["b"], out["j"]


def inner_fold(ic: dict) -> np.ndarray:
    """(3 n_r_ic_max, n_r_ic_max): inner values, d/dr and d2/dr2 from the inner
    even-Chebyshev coefficients."""
    Ci, D1i, D2i = ic["CT"], ic["D1"], ic["D2"]
    return np.concatenate([Ci, D1i @ Ci, D2i @ Ci])


def row_operators(p, rad: dict, ic: dict, inv: np.ndarray, k4: int):
    """Row-operator matrices MT[l][k][r] (k < k4 padded, r padded to 64) of one
    coupled system: outer (values, d/dr, d2/dr2 on n_r_max levels) and inner
    (the same on n_r_ic_max levels), each acting on the n_r_tot right-hand side."""
    N, Nic = p.n_r_max, p.n_r_ic_max
    CT, D1, D2 = rad["CT"], rad["D1"], rad["D2"]
    Ci, D1i, D2i = ic["CT"], ic["D1"], ic["D2"]
    op_oc = np.concatenate([CT, D1 @ CT, D2 @ CT])          # (3N, N) on outer coefficients
    op_ic = np.concatenate([Ci, D1i @ Ci, D2i @ Ci])        # (3Nic, Nic) on inner ones
    mats = []
    for op, part in ((op_oc, inv[:, :N, :]), (op_ic, inv[:, N:N + Nic, :])):
        full = np.einsum("rc,lck->lkr", op, part)           # (l, NT, R) = (op @ inv_part)^T
        R = op.shape[0]
        Rp = -(-R // 64) * 64
        mt = np.zeros((p.l_max + 1, k4, Rp))
        mt[:, :full.shape[1], :R] = full
        mats.append((mt, R, Rp))
    return mats


def rotation_constants(p) -> dict:
    """Coefficients of the inner core's rotation (reference
    ``pre_calculations.py:56-92``, Boussinesq): omega_ic = c_z10_omega z10(r_icb),
    old state c_dt_z10 z10(r_icb), torque term c_lorentz * torque, and the
    torque's quadrature factor LFfac 2 pi / n_phi over one sector of the grid."""
    or2_icb = (1.0 - p.radratio) ** 2 / p.radratio ** 2
    return dict(c_z10_omega=0.5 * math.sqrt(3.0 / math.pi) * or2_icb,
                c_dt_z10=0.2 * p.radratio / (1.0 - p.radratio),
                c_lorentz=0.25 * math.sqrt(3.0 / math.pi) * or2_icb,
                fac_lt=p.LFfac * 2.0 * math.pi / p.n_phi)


def z10_inverse(p, rad: dict, wimp_lin0: float, c_dt_z10: float) -> np.ndarray:
    """(N, N) inverse of the z(l=1, m=0) system of a rotating inner core with the
    row scaling folded in (reference ``update_z.py:244-298``, time-integrated
    rotation): the l = 1 toroidal diffusion rows, z = 0 at the CMB, and at the ICB
    the angular-momentum balance c_dt_z10 z + wimp_lin0 (2 z / r - dz/dr)."""
    a = z10_matrix(p, rad, wimp_lin0, c_dt_z10)
    fac = 1.0 / np.abs(a).max(axis=1)
    return np.linalg.inv(fac[:, None] * a) * fac[None, :]


def z10_affine(p, rad: dict, c_dt_z10: float):
    """(A0, A1) with the z(l=1, m=0) matrix A0 + w A1."""
    a0 = z10_matrix(p, rad, 0.0, c_dt_z10)
    return a0, z10_matrix(p, rad, 1.0, c_dt_z10) - a0


def z10_matrix(p, rad: dict, wimp_lin0: float, c_dt_z10: float) -> np.ndarray:
    R, D, D2 = rad["T"], rad["T1"], rad["T2"]
    N, rn = p.n_r_max, rad["rnorm"]
    or1, o2 = rad["or1"], rad["or2"][:, None]
    a = rn * 2.0 * o2 * (R - wimp_lin0 * (D2 - 2.0 * o2 * R))
    a[0] = rn * R[0]
    a[N - 1] = rn * (c_dt_z10 * R[N - 1] + wimp_lin0 * (2.0 * or1[N - 1] * R[N - 1] - D[N - 1]))
    if p.n_cheb_max < N:
        a[(0, N - 1), p.n_cheb_max:] = 0.0
    a[:, (0, N - 1)] *= rad.get("bfac", 0.5)
    return a


def z10_row_operator(p, rad: dict, inv: np.ndarray, k4: int, rpad: int) -> np.ndarray:
    """MT[k][r] (k < k4, r < rpad) of the z(l=1, m=0) solve: values, d/dr and
    d2/dr2 on the N levels from the right-hand side."""
    N, CT = p.n_r_max, rad["CT"]
    op = np.concatenate([CT, rad["D1"] @ CT, rad["D2"] @ CT]) @ inv     # (3N, N)
    mt = np.zeros((k4, rpad))
    mt[:N, :3 * N] = op.T
    return mt
