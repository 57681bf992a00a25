"""Host set-up of the finite-difference time step (double-curl poloidal equation).

With ``radial_scheme = "FD"`` the reference eliminates the pressure for
l >= 1 (``params.py:143``): w obeys a fourth-order equation solved as one N x N
system per degree, and only the l = 0 pressure is solved (from the hydrostatic
balance).  Everything here runs once per run on the CPU in FP64; the device
turns the assembled operators into row operators on every dt change
(``mh_affine_f64`` + ``mh_gj_inverse`` + ``mh_gemm_f64``).

* ``radial`` -- the resolution's radial data on the stretched grid
  (``finite_differences.py:62-351``): identity collocation (rMat = I), dense
  derivative matrices from the Fornberg stencils, rnorm = 1, no DCT, Simpson
  weights for radial integrals (``integration.py:75-138``).
* ``affine_operators`` -- every implicit operator as A0 + wimp A1, per degree:
  entropy (``update_s.py:121-143``), toroidal (``update_z.py:160-190``),
  double-curl poloidal (``update_w_doublecurl.py:153-223``), insulating
  poloidal and toroidal magnetic (``update_b.py:160-230``); the boundary rows
  sit in A0.  Degrees without an equation (l = 0 of Z, W, B, J) get the
  identity; their row operators are zeroed after the fold.
* ``p0_inverse`` -- the l = 0 pressure matrix (``update_wp.py:95-158``): rows
  d/dr - beta, the last row a first-order backward difference cut to the
  Fortran band (order/2 below the diagonal), p(r_cmb) = 0 in row 0.
* ``conduction_state`` -- (s0, p0) of the 2N conduction system
  (``init_fields.py:108-182,183-311``; the trapezoidal integral row of
  anelastic runs with viscous heating).
"""

from __future__ import annotations

import math

import numpy as np

from . import fd


def dense_derivatives(r: np.ndarray, order: int, order_boundary: int):
    """(D1, D2, D3, D4) dense, rows as ``finite_differences.py:279-331``."""
    n = len(r)
    hw = max(order // 2 + 1, order_boundary + 4)
    bands = fd.derivative_bands(r, order, order_boundary, hw)
    out = []
    for k in (1, 2, 3, 4):
        B = bands[k]
        D = np.zeros((n, n))
        for d in range(-hw, hw + 1):
            i = np.arange(max(0, -d), min(n, n - d))
            D[i, i + d] = B[i, hw + d]
        out.append(D)
    return tuple(out)


def simpson_weights(r: np.ndarray) -> np.ndarray:
    """q with integral f dr = f . q: three-point panels on the non-uniform grid
    (odd n), or for even n the mean of the two covers that close with one
    trapezoid at either end; r decreases, so the sum is negated."""
    n = len(r)
    q = np.zeros(n)

    def add_panel(i, s):
        h1, h2 = r[i] - r[i - 1], r[i + 1] - r[i]
        c = s * (h1 + h2) / 6.0
        q[i - 1] += c * (2.0 * h1 - h2) / h1
        q[i] += c * (h1 + h2) ** 2 / (h1 * h2)
        q[i + 1] += c * (2.0 * h2 - h1) / h2

    if n % 2:
        for i in range(1, n - 1, 2):
            add_panel(i, -1.0)
        return q
    t0, t1 = 0.5 * (r[1] - r[0]), 0.5 * (r[n - 1] - r[n - 2])
    q[0] -= 0.5 * t0
    q[1] -= 0.5 * t0
    q[n - 2] -= 0.5 * t1
    q[n - 1] -= 0.5 * t1
    for i in range(1, n - 1):
        add_panel(i, -0.5)
    return q


def radial(p) -> dict:
    """Radial data of a finite-difference resolution (same keys as the
    Chebyshev set-up; T = C = CT = I, T1..T3 = D1..D3)."""
    N = p.n_r_max
    r_cmb = 1.0 / (1.0 - p.radratio)
    r_icb = r_cmb - 1.0
    r = fd.stretched_grid(N, r_icb, r_cmb, p.fd_stretch, p.fd_ratio)
    D1, D2, D3, D4 = dense_derivatives(r, p.fd_order, p.fd_order_bound)
    eye = np.eye(N)
    dr = r[:-1] - r[1:]
    delxr2 = np.empty(N)
    delxr2[0], delxr2[-1] = dr[0] ** 2, dr[-1] ** 2
    delxr2[1:-1] = np.minimum(dr[:-1], dr[1:]) ** 2
    delxh2 = r ** 2 / float(p.l_max * (p.l_max + 1))
    fac = (0.5 * (1.0 + p.opm)) ** 2
    return dict(r=r, x=None, r_cmb=r_cmb, r_icb=r_icb, rnorm=1.0, T=eye, T1=D1, T2=D2, T3=D3,
                D1=D1, D2=D2, D3=D3, D4=D4, C=eye, CT=eye, q=simpson_weights(r),
                or1=1.0 / r, or2=r ** -2.0, or3=r ** -3.0, or4=r ** -4.0, r2=r * r,
                rgrav=r / r_cmb, delxr2=delxr2, delxh2=delxh2, valri2=fac / delxr2,
                valhi2=fac / delxh2, fd=True, bfac=1.0)


def boussinesq_profile(rad) -> dict:
    n = len(rad["r"])
    z, o = np.zeros(n), np.ones(n)
    return dict(beta=z, dbeta=z, ddbeta=z, dLtemp0=z, rho0=o, orho1=o, visc=o, dLvisc=z,
                ddLvisc=z, kappa=o, dLkappa=z, rgrav=rad["rgrav"], temp0=o, alpha0=o, ogrun=z,
                ViscHeatFac=0.0, ThExpNb=1.0)


def affine_operators(p, rad, prof) -> dict:
    """{"szbj": (A0, A1) of shape (nsys, l_max+1, N, N) for S, Z (B, J),
    "w": (A0, A1) of shape (l_max+1, N, N)}."""
    N, lp1 = p.n_r_max, p.l_max + 1
    D1, D2, D3, D4 = rad["D1"], rad["D2"], rad["D3"], rad["D4"]
    I = np.eye(N)
    or1, or2 = rad["or1"], rad["or2"]
    col = lambda v: v[:, None]
    be, dbe, ddbe = prof["beta"], prof["dbeta"], prof["ddbeta"]
    orho, visc, dv, ddv = prof["orho1"], prof["visc"], prof["dLvisc"], prof["ddLvisc"]
    l = np.arange(lp1, dtype=np.float64)
    dL = (l * (l + 1.0))[:, None, None]
    nsys = 4 if p.l_mag else 2
    A0 = np.zeros((nsys, lp1, N, N))
    A1 = np.zeros((nsys, lp1, N, N))
    # entropy (every degree): I - w opr kappa (D2 + (beta + dLtemp0 + 2/r + dLkappa) D1 - dL/r^2)
    d1s = col(be + prof["dLtemp0"] + 2.0 * or1 + prof["dLkappa"])
    A0[0] = I
    A1[0] = -p.opr * col(prof["kappa"]) * (D2 + d1s * D1 - dL * col(or2) * I)
    # toroidal: dL/r^2 (I - w visc (D2 + (dLvisc - beta) D1 - (...) I))
    zc = col(dv * be + 2.0 * dv * or1 + dbe + 2.0 * be * or1)
    A0[1] = dL * col(or2) * I
    A1[1] = -dL * col(or2 * visc) * (D2 + col(dv - be) * D1 - (zc + dL * col(or2)) * I)
    if p.l_mag:
        A0[2] = A0[3] = dL * col(or2) * I
        A1[2] = A1[3] = -p.opm * dL * col(or2) * (D2 - dL * col(or2) * I)
    A1[:, :, 0] = 0.0          # boundary rows carry no implicit part
    A1[:, :, N - 1] = 0.0

    def wall(k, i, kv):   # no-slip (z = 0) or stress-free (dz/dr - (2/r + beta) z = 0)
        A0[k, 1:, i] = I[i] if kv != 1 else D1[i] - (2.0 * or1[i] + be[i]) * I[i]

    wall(1, 0, p.ktopv)
    wall(1, N - 1, p.kbotv)
    if p.l_mag:
        for li in range(1, lp1):
            A0[2, li, 0] = D1[0] + li * or1[0] * I[0]
            A0[2, li, N - 1] = D1[N - 1] - (li + 1) * or1[N - 1] * I[N - 1]
        A0[3, 1:, 0], A0[3, 1:, N - 1] = I[0], I[N - 1]
    # double-curl poloidal
    c3 = 2.0 * (dv - be)
    c2 = ddv - 2.0 * dbe + dv ** 2 + be ** 2 - 3.0 * dv * be - 2.0 * or1 * (dv + be)
    c1 = (-ddbe - dbe * (2.0 * dv - be + 2.0 * or1) - ddv * (be + 2.0 * or1)
          + be ** 2 * (dv + 2.0 * or1) - be * (dv ** 2 - 2.0 * or2) - 2.0 * dv * or1 * (dv - or1))
    c1l = 2.0 * (2.0 * or1 + be - dv) * or2
    c0 = 2.0 * dbe + ddv + dv ** 2 - (2.0 / 3.0) * be ** 2 + dv * be + 2.0 * or1 * (
        2.0 * dv - be - 3.0 * or1)
    W0 = -dL * col(or2 * orho) * (D2 - col(be) * D1 - dL * col(or2) * I)
    W1 = dL * col(orho * visc * or2) * (
        D4 + col(c3) * D3 + (col(c2) - 2.0 * dL * col(or2)) * D2 + (col(c1) + dL * col(c1l)) * D1
        + dL * col(or2) * (col(c0) + dL * col(or2)) * I)
    for i in (0, 1, N - 2, N - 1):
        W1[:, i] = 0.0
    W0[:, 0], W0[:, N - 1] = I[0], I[N - 1]

    def wall_w(i, j, kv):
        W0[:, i] = D2[j] - (2.0 * or1[j] + be[j]) * D1[j] if kv == 1 else D1[j]

    wall_w(1, 0, p.ktopv)
    wall_w(N - 2, N - 1, p.kbotv)
    # degrees without an equation
    for k in range(1, nsys):
        A0[k, 0], A1[k, 0] = I, 0.0
    W0[0], W1[0] = I, 0.0
    return {"szbj": (A0, A1), "w": (W0, W1)}


def p0_inverse(p, rad, prof) -> np.ndarray:
    N, r, be, D1 = p.n_r_max, rad["r"], prof["beta"], rad["D1"]
    d = D1 - np.diag(be)
    delr = r[N - 1] - r[N - 2]
    d[N - 1] = 0.0
    d[N - 1, N - 1] = 1.0 / delr - be[N - 1]
    d[N - 1, N - 2] = -1.0 / delr
    kl = p.fd_order // 2
    last = D1[N - 1].copy()
    keep = np.arange(N) >= N - 1 - kl
    d[N - 1, keep & (np.arange(N) < N - 2)] = last[keep & (np.arange(N) < N - 2)]
    d[0] = 0.0
    d[0, 0] = 1.0
    return np.linalg.inv(d)


def conduction_state(p, rad, prof):
    """(s0, p0) on the grid: opr kappa (s'' + (beta + dLtemp0 + 2/r + dLkappa) s') = 0
    with s = 0 outside and sqrt(4 pi) inside; p' - beta p = rho0 BuoFac g s with
    p(r_cmb) = 0 or, with viscous heating, the trapezoidal integral row."""
    N, r = p.n_r_max, rad["r"]
    D1, D2, I = rad["D1"], rad["D2"], np.eye(N)
    M = np.zeros((2 * N, 2 * N))
    d1 = prof["beta"] + prof["dLtemp0"] + 2.0 * rad["or1"] + prof["dLkappa"]
    M[:N, :N] = p.opr * prof["kappa"][:, None] * (D2 + d1[:, None] * D1)
    M[N:, :N] = -np.diag(prof["rho0"] * p.BuoFac * prof["rgrav"])
    M[N:, N:] = D1 - np.diag(prof["beta"])
    M[0] = 0.0
    M[0, 0] = 1.0
    M[N - 1] = 0.0
    M[N - 1, N - 1] = 1.0
    M[N] = 0.0
    vt = prof["ViscHeatFac"] * prof["ThExpNb"]
    if p.l_anel and vt != 0.0:
        h = np.empty(N)
        h[1:N - 1] = r[2:] - r[:-2]
        h[0], h[N - 1] = r[1] - r[0], r[N - 1] - r[N - 2]
        r2 = r * r
        M[N, :N] = -0.5 * prof["ThExpNb"] * prof["alpha0"] * prof["temp0"] * prof["rho0"] * r2 * h
        M[N, N:] = 0.5 * vt * prof["ogrun"] * prof["alpha0"] * r2 * h
    else:
        M[N, N] = 1.0
    fac = 1.0 / np.abs(M).max(axis=1)
    rhs = np.zeros(2 * N)
    rhs[N - 1] = math.sqrt(4.0 * math.pi)
    x = np.linalg.solve(fac[:, None] * M, fac * rhs)
    return x[:N], x[N:]


def to_band(A: np.ndarray):
    """(abd, k): LINPACK storage abd[..., k + k + i - j, j] = A[..., i, j] with
    2k + k + 1 rows (the first k left for fill-in) of a batch of matrices whose
    nonzeros lie within k diagonals of the main one, k the smallest such."""
    n = A.shape[-1]
    nz = np.abs(A.reshape(-1, n, n)).max(axis=0) != 0
    i, j = np.nonzero(nz)
    k = int(np.abs(i - j).max()) if i.size else 0
    abd = np.zeros(A.shape[:-2] + (3 * k + 1, n))
    for d in range(-k, k + 1):       # d = j - i
        rows = np.arange(max(0, -d), min(n, n - d))
        abd[..., 2 * k - d, rows + d] = A[..., rows, rows + d]
    return abd, k


def band_operators(p, rad, prof) -> dict:
    """The affine operators of ``affine_operators`` in LINPACK band storage for
    the device's banded factorisation: {"szbj": (B0, B1, k), "w": (W0, W1, k)}."""
    ops = affine_operators(p, rad, prof)
    out = {}
    for key in ("szbj", "w"):
        a0, a1 = ops[key]
        b, k = to_band(np.stack([a0, a1]))
        out[key] = (b[0], b[1], k)
    return out
