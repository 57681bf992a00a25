"""Radial discretisation (oracle): Chebyshev grid, DCT-I, derivative matrices.

Restates ``chebyshev.py:20-107`` (Chebyshev-Gauss-Lobatto grid on
[r_icb, r_cmb], rMat..d3rMat by the T_n recursion with the linear-map chain
rule), ``cosine_transform.py:18-45`` (self-inverse DCT-I, scaled 1/sqrt(2N)),
``radial_derivatives.py:71-106`` (D_k = d^k rMat · rMat^-1 with n_cheb_max
truncation), ``integration.py:18-72`` (Clenshaw-Curtis rInt_R) and the
Boussinesq part of ``radial_functions.py:18-69,189-202``.

Pinned by ``tests/test_oracle_radial.py``: rMat against numpy's Chebyshev
T_n, D1..D3 exact on polynomials of degree < N, costf against
``scipy.fft.dct(type=1)`` times 1/sqrt(2N) and costf∘costf = I, rInt_R exact
on polynomials.
"""

from __future__ import annotations

import math

import numpy as np

from .params import Params


def cheb_grid(p: Params):
    """Grid r[0]=r_cmb .. r[N-1]=r_icb (``chebyshev.py:20-45``)."""
    r_cmb = 1.0 / (1.0 - p.radratio)
    r_icb = r_cmb - 1.0
    n = p.n_r_max - 1
    x = np.cos(math.pi * np.arange(n + 1) / n)
    r = 0.5 * (r_cmb - r_icb) * x + 0.5 * (r_icb + r_cmb)
    return r, x, r_cmb, r_icb


def cheb_mats(p: Params, x=None, r_cmb=None, r_icb=None):
    """rMat, drMat, d2rMat, d3rMat (``chebyshev.py:53-107``)."""
    if x is None:
        _, x, r_cmb, r_icb = cheb_grid(p)
    n = p.n_r_max
    drx = 2.0 / (r_cmb - r_icb)
    R = np.zeros((n, n)); D = np.zeros((n, n)); D2 = np.zeros((n, n)); D3 = np.zeros((n, n))
    R[:, 0] = 1.0
    R[:, 1] = x
    D[:, 1] = drx
    for c in range(2, n):
        R[:, c] = 2 * x * R[:, c - 1] - R[:, c - 2]
        D[:, c] = 2 * drx * R[:, c - 1] + 2 * x * D[:, c - 1] - D[:, c - 2]
        D2[:, c] = 4 * drx * D[:, c - 1] + 2 * x * D2[:, c - 1] - D2[:, c - 2]
        D3[:, c] = 6 * drx * D2[:, c - 1] + 2 * x * D3[:, c - 1] - D3[:, c - 2]
    return R, D, D2, D3


def costf(f: np.ndarray) -> np.ndarray:
    """Self-inverse DCT-I along the last axis (``cosine_transform.py:18-45``)."""
    n = f.shape[-1]
    N = n - 1
    ext = np.concatenate([f, f[..., 1:-1][..., ::-1]], axis=-1)
    F = np.fft.fft(ext, axis=-1)[..., :n]
    if not np.iscomplexobj(f):
        F = F.real
    return F / math.sqrt(2.0 * N)


def costf_matrix(n: int) -> np.ndarray:
    """Dense matrix C with costf(f) = f @ C.T for real f (same transform)."""
    return costf(np.eye(n)).T


def diff_mats(p: Params, mats=None):
    """D1, D2, D3 acting on grid values (``radial_derivatives.py:71-106``)."""
    R, D, D2, D3 = mats if mats is not None else cheb_mats(p)
    Rinv = np.linalg.inv(R)
    if p.n_cheb_max < p.n_r_max:
        Rinv = Rinv.copy()
        Rinv[p.n_cheb_max:, :] = 0.0
    return D @ Rinv, D2 @ Rinv, D3 @ Rinv


def cheb_int_weights(n: int) -> np.ndarray:
    """Chebyshev integration weights (``integration.py:18-38``)."""
    w = np.zeros(n)
    w[0] = 1.0
    for k in range(2, n, 2):
        k1 = k + 1
        w[k] = -1.0 / (k1 * (k1 - 2))
    return w


def rint_weights(p: Params, r_cmb: float, r_icb: float) -> np.ndarray:
    """Vector q with rInt_R(f) = f @ q (``integration.py:45-72`` is linear in f)."""
    n = p.n_r_max
    drx = 2.0 / (r_cmb - r_icb)
    rnorm = math.sqrt(2.0 / (n - 1))
    w = cheb_int_weights(n)
    C = costf_matrix(n)             # a = C @ g
    bf = np.ones(n); bf[0] = 0.5; bf[-1] = 0.5
    # result = 2 rnorm sum_k w_k bf_k (C @ (f/drx))_k
    return 2.0 * rnorm * (C.T @ (w * bf)) / drx


def ic_grid(p: Params, r_icb: float) -> dict:
    """Inner-core grid and even Chebyshev polynomials (``radial_functions.py:209-306``,
    ``radial_derivatives.py:182-195``): the upper half of the 2 n_r_ic_max - 1
    point Gauss-Lobatto grid on [-r_icb, r_icb] (r_ic[0] = r_icb .. r_ic[-1] = 0),
    cheb[n, k] = T_2n(y_k) with its first and second r-derivatives by the
    recursion over T_2n and T_2n+1 (map factor 1/r_icb), the norm
    sqrt(2/(n_r_ic_max - 1)) and the grid-to-grid derivative matrices
    D_k = cheb^-1 d^k cheb (applied as f @ D_k)."""
    N = p.n_r_ic_max
    n_tot = 2 * N - 1
    y = [math.cos(math.pi * k / (n_tot - 1)) for k in range(n_tot)]
    r_ic = np.zeros(N)
    r_ic[:N - 1] = [r_icb * y[k] for k in range(N - 1)]
    O_r_ic = np.zeros(N)
    O_r_ic[:N - 1] = 1.0 / r_ic[:N - 1]
    mf = 1.0 / r_icb
    cheb = np.zeros((N, N)); dcheb = np.zeros((N, N)); d2cheb = np.zeros((N, N))
    for k in range(N):
        yk = y[k]
        cheb[0, k] = 1.0
        lc, ldc, ld2c = yk, mf, 0.0   # T_1 and its derivatives
        for n in range(1, N):
            cheb[n, k] = 2.0 * yk * lc - cheb[n - 1, k]
            dcheb[n, k] = 2.0 * mf * lc + 2.0 * yk * ldc - dcheb[n - 1, k]
            d2cheb[n, k] = 4.0 * mf * ldc + 2.0 * yk * ld2c - d2cheb[n - 1, k]
            lc, ldc, ld2c = (2.0 * yk * cheb[n, k] - lc,
                             2.0 * mf * cheb[n, k] + 2.0 * yk * dcheb[n, k] - ldc,
                             4.0 * mf * dcheb[n, k] + 2.0 * yk * d2cheb[n, k] - ld2c)
    inv = np.linalg.inv(cheb)
    return dict(r_ic=r_ic, O_r_ic=O_r_ic, cheb=cheb, dcheb=dcheb, d2cheb=d2cheb,
                cnorm=math.sqrt(2.0 / (N - 1)), D1=inv @ dcheb, D2=inv @ d2cheb)


def rint(f: np.ndarray, q: np.ndarray):
    """Clenshaw-Curtis integral over r of f(..., N) (``integration.py:45-72``)."""
    return f @ q


def simps_weights(r: np.ndarray) -> np.ndarray:
    """Vector q with simps(f, r) = f @ q (``integration.py:75-122``, linear in f):
    composite three-point panels on the non-uniform grid for odd N; for even N
    the mean of (trapezoid on the first interval + panels centred on the even
    points) and (panels centred on the odd points + trapezoid on the last
    interval); negated because r decreases."""
    n = r.shape[0]
    q = np.zeros(n)

    def panel(i, s):
        h2, h1 = r[i + 1] - r[i], r[i] - r[i - 1]
        c = s * (h1 + h2) / 6.0
        q[i - 1] += c * (2.0 * h1 - h2) / h1
        q[i] += c * (h1 + h2) ** 2 / (h1 * h2)
        q[i + 1] += c * (2.0 * h2 - h1) / h2

    if n % 2 == 1:
        for i in range(1, n - 1, 2):
            panel(i, -1.0)
        return q
    s = -0.5
    q[0] += s * 0.5 * (r[1] - r[0]); q[1] += s * 0.5 * (r[1] - r[0])
    for i in range(2, n - 1, 2):
        panel(i, s)
    q[n - 1] += s * 0.5 * (r[n - 1] - r[n - 2]); q[n - 2] += s * 0.5 * (r[n - 1] - r[n - 2])
    for i in range(1, n - 1, 2):
        panel(i, s)
    return q


def fd_mats(p: Params, r: np.ndarray):
    """(rMat = I, drMat, d2rMat, d3rMat, d4rMat) of the finite-difference scheme
    (``finite_differences.py:279-331``), dense."""
    from . import fd
    c = fd.fd_coeffs(r, p.fd_order, p.fd_order_bound)
    return fd.fd_der_mats(r, c, p.fd_order, p.fd_order_bound)


def radial_functions(p: Params):
    """Boussinesq radial profiles and CFL spacings (``radial_functions.py:18-69,189-202``)
    on the Chebyshev grid, or on the stretched finite-difference grid
    (``finite_differences.py:62-127,342-351``: rnorm = boundary_fac = 1)."""
    if p.l_finite_diff:
        from . import fd
        r_cmb = 1.0 / (1.0 - p.radratio)
        r_icb = r_cmb - 1.0
        r = fd.fd_grid(p.n_r_max, r_icb, r_cmb, p.fd_stretch, p.fd_ratio)
        x, rnorm, bfac = None, 1.0, 1.0
    else:
        r, x, r_cmb, r_icb = cheb_grid(p)
        rnorm, bfac = math.sqrt(2.0 / (p.n_r_max - 1)), 0.5
    or1 = 1.0 / r
    dr = r[:-1] - r[1:]
    delxr2 = np.empty(p.n_r_max)
    delxr2[0] = dr[0] ** 2
    delxr2[-1] = dr[-1] ** 2
    delxr2[1:-1] = np.minimum(dr[:-1], dr[1:]) ** 2
    return dict(r=r, x=x, r_cmb=r_cmb, r_icb=r_icb, or1=or1, or2=or1 ** 2, or3=or1 ** 3,
                or4=or1 ** 4, rgrav=r / r_cmb,
                delxh2=r ** 2 / float(p.l_max * (p.l_max + 1)), delxr2=delxr2,
                rnorm=rnorm, bfac=bfac)
