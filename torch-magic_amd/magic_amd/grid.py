"""Host-side set-up of one resolution: index maps, quadrature, radial operators.

Everything here runs once per (l_max, n_r_max, radratio) on the CPU in FP64
and is uploaded to the device by :mod:`magic_amd.device`; nothing in this
module is on the per-step path.

* lm maps — standard m-major order (reference ``blocking.py:29-35``).
* Gauss-Legendre nodes/weights — Golub-Welsch eigenvalues refined by Newton
  steps on P_n (same nodes as the reference ``horizontal_data.py:16-52``).
* Chebyshev-Gauss-Lobatto radial grid with T_n, dT_n/dr, d2, d3 collocation
  matrices (reference ``chebyshev.py:20-107``), the self-inverse DCT-I as a
  matrix (reference ``cosine_transform.py:18-45``) and the derivative
  operators D_k = (d^k T)(T)^-1 with n_cheb_max truncation (reference
  ``radial_derivatives.py:71-106``), Clenshaw-Curtis radial weights (reference
  ``integration.py:18-72``).
"""

from __future__ import annotations

import math

import numpy as np

from .params import Params


def lm_maps(p: Params) -> dict:
    """m-major lm order over the orders m = mi * minc (mode index mi < n_m):
    lm_start[mi] is the row of (l = m, m); l / m: degree and order of each row;
    l_rows[l_rows_off[l]:l_rows_off[l + 1]]: the rows of degree l."""
    l_max, minc = p.l_max, p.minc
    orders = minc * np.arange(p.n_m)
    lm_start = np.zeros(p.n_m + 1, np.int64)
    lm_start[1:] = np.cumsum(l_max + 1 - orders)
    m_of = np.repeat(orders, l_max + 1 - orders)
    l_of = np.concatenate([np.arange(m, l_max + 1) for m in orders])
    # rows of each degree l: lm(l, m) for the orders m <= l
    l_rows = np.concatenate([lm_start[:l // minc + 1] + (l - orders[:l // minc + 1])
                             for l in range(l_max + 1)])
    l_rows_off = np.concatenate([[0], np.cumsum(np.arange(l_max + 1) // minc + 1)])
    return dict(lm_start=lm_start, l=l_of, m=m_of, l_rows=l_rows, l_rows_off=l_rows_off)


def gauss_legendre(n: int):
    """Nodes x_j = cos(theta_j), theta ascending (north first), and weights."""
    k = np.arange(1, n)
    beta = k / np.sqrt(4.0 * k * k - 1.0)
    x = np.linalg.eigvalsh(np.diag(beta, 1) + np.diag(beta, -1))
    for _ in range(3):  # Newton refinement on P_n(x)
        p0, p1 = np.ones_like(x), x.copy()
        for j in range(2, n + 1):
            p0, p1 = p1, ((2 * j - 1) * x * p1 - (j - 1) * p0) / j
        dp = n * (x * p1 - p0) / (x * x - 1.0)
        x = x - p1 / dp
    p0, p1 = np.ones_like(x), x.copy()
    for j in range(2, n + 1):
        p0, p1 = p1, ((2 * j - 1) * x * p1 - (j - 1) * p0) / j
    dp = n * (x * p1 - p0) / (x * x - 1.0)
    w = 2.0 / ((1.0 - x * x) * dp * dp)
    order = np.argsort(-x)
    x, w = x[order], w[order]
    # exact symmetry of the pairs
    h = n // 2
    x[n - h:] = -x[:h][::-1]
    w[n - h:] = w[:h][::-1]
    return x, w


def theta_grid(p: Params) -> dict:
    x, w = gauss_legendre(p.n_theta)
    H = p.n_theta // 2
    theta = np.arccos(x)
    tn = theta[:H]
    sin_n = np.sin(tn)
    cos_n = np.cos(tn)
    # grid rows interleave the northern node j and its southern mirror
    sin_grid = np.repeat(sin_n, 2)
    cos_grid = np.empty(p.n_theta)
    cos_grid[0::2] = cos_n
    cos_grid[1::2] = -cos_n
    return dict(H=H, cos_n=cos_n, sin_n=sin_n, w_n=w[:H], theta_n=tn, sin_grid=sin_grid,
                cos_grid=cos_grid, ost2_grid=1.0 / sin_grid ** 2, gauss_grid=np.repeat(w[:H], 2))


def lm_coefficients(p: Params, maps: dict) -> dict:
    l = maps["l"].astype(np.float64)
    m = maps["m"].astype(np.float64)

    def clm(ll):
        out = np.zeros_like(ll)
        ok = ll > 0
        out[ok] = np.sqrt((ll[ok] + m[ok]) * (ll[ok] - m[ok]) / ((2 * ll[ok] - 1) * (2 * ll[ok] + 1)))
        return out

    c0, c1 = clm(l), clm(l + 1)
    return dict(dLh=l * (l + 1), dTh2A=(l + 2) * c1, dTh2S=(l - 1) * c0,
                dTh3A=l * (l + 2) * c1, dTh3S=(l - 1) * (l + 1) * c0,
                dTh4A=l * c1 * (l + 1) * (l + 2), dTh4S=(l + 1) * c0 * (l - 1) * l)


def radial(p: Params) -> dict:
    if p.l_finite_diff:   # stretched finite-difference grid (magic_amd.fdstep)
        from . import fdstep
        return fdstep.radial(p)
    N = p.n_r_max
    r_cmb = 1.0 / (1.0 - p.radratio)
    r_icb = r_cmb - 1.0
    x = np.cos(math.pi * np.arange(N) / (N - 1))
    r = 0.5 * (r_cmb - r_icb) * x + 0.5 * (r_cmb + r_icb)
    drx = 2.0 / (r_cmb - r_icb)
    # T_n(x_i) and r-derivatives by the three-term recursion (linear map)
    T = np.zeros((N, N)); T1 = np.zeros((N, N)); T2 = np.zeros((N, N)); T3 = np.zeros((N, N))
    T[:, 0] = 1.0
    T[:, 1] = x
    T1[:, 1] = drx
    for n in range(2, N):
        T[:, n] = 2 * x * T[:, n - 1] - T[:, n - 2]
        T1[:, n] = 2 * drx * T[:, n - 1] + 2 * x * T1[:, n - 1] - T1[:, n - 2]
        T2[:, n] = 4 * drx * T1[:, n - 1] + 2 * x * T2[:, n - 1] - T2[:, n - 2]
        T3[:, n] = 6 * drx * T2[:, n - 1] + 2 * x * T3[:, n - 1] - T3[:, n - 2]
    Tinv = np.linalg.inv(T)
    if p.n_cheb_max < N:
        Tinv[p.n_cheb_max:] = 0.0
    D1, D2, D3 = T1 @ Tinv, T2 @ Tinv, T3 @ Tinv
    # DCT-I, self-inverse with the 1/sqrt(2(N-1)) scale: C[k, j]
    kk = np.arange(N)
    C = np.cos(math.pi * np.outer(kk, kk) / (N - 1))
    C[:, 1:N - 1] *= 2.0
    C /= math.sqrt(2.0 * (N - 1))
    trunc = np.ones(N)
    trunc[p.n_cheb_max:] = 0.0
    CT = C * trunc[None, :]
    # Clenshaw-Curtis: integral f dr = f . q
    cheb_int = np.zeros(N)
    cheb_int[0] = 1.0
    for k in range(2, N, 2):
        cheb_int[k] = -1.0 / ((k + 1) * (k - 1))
    bf = np.ones(N); bf[0] = bf[-1] = 0.5
    rnorm = math.sqrt(2.0 / (N - 1))
    q = 2.0 * rnorm * (C.T @ (cheb_int * bf)) / drx
    dr = r[:-1] - r[1:]
    delxr2 = np.empty(N)
    delxr2[0], delxr2[-1] = dr[0] ** 2, dr[-1] ** 2
    delxr2[1:-1] = np.minimum(dr[:-1], dr[1:]) ** 2
    delxh2 = r ** 2 / float(p.l_max * (p.l_max + 1))
    fac = (0.5 * (1.0 + p.opm)) ** 2
    return dict(r=r, x=x, r_cmb=r_cmb, r_icb=r_icb, rnorm=rnorm, T=T, T1=T1, T2=T2, T3=T3,
                D1=D1, D2=D2, D3=D3, C=C, CT=CT, q=q, or1=1.0 / r, or2=r ** -2.0,
                or3=r ** -3.0, or4=r ** -4.0, r2=r * r, rgrav=r / r_cmb,
                delxr2=delxr2, delxh2=delxh2, valri2=fac / delxr2, valhi2=fac / delxh2)


def inner_core(p: Params, r_icb: float) -> dict:
    """Inner-core radial grid of a conducting inner core (reference
    ``radial_functions.py:209-306``, ``radial_derivatives.py:182-195``): the
    upper half r_ic[k] = r_icb y_k (k < n_r_ic_max - 1, then the centre r = 0)
    of the 2 n_r_ic_max - 1 point Gauss-Lobatto grid y_k = cos(pi k / (2
    n_r_ic_max - 2)) on [-r_icb, r_icb]; the even Chebyshev polynomials
    cheb[n, k] = T_2n(y_k) and their r-derivatives (map factor 1/r_icb, the
    three-term recursion run over T_2n and T_2n+1); the grid-to-grid derivative
    operators as column operators (df = D1 f); the DCT-I of n_r_ic_max points
    with the n_cheb_ic_max truncation (values = CT coefficients)."""
    n = p.n_r_ic_max
    n_tot = 2 * n - 1
    y = np.cos(math.pi * np.arange(n_tot) / (n_tot - 1))
    r_ic = np.zeros(n)
    r_ic[:n - 1] = r_icb * y[:n - 1]
    O_r_ic = np.zeros(n)
    O_r_ic[:n - 1] = 1.0 / r_ic[:n - 1]
    mf = 1.0 / r_icb
    T = np.zeros((n, n)); T1 = np.zeros((n, n)); T2 = np.zeros((n, n))
    yk = y[:n]
    T[0] = 1.0
    odd, odd1, odd2 = yk.copy(), np.full(n, mf), np.zeros(n)   # T_1 and its derivatives
    for j in range(1, n):
        T[j] = 2.0 * yk * odd - T[j - 1]
        T1[j] = 2.0 * mf * odd + 2.0 * yk * odd1 - T1[j - 1]
        T2[j] = 4.0 * mf * odd1 + 2.0 * yk * odd2 - T2[j - 1]
        odd, odd1, odd2 = (2.0 * yk * T[j] - odd, 2.0 * mf * T[j] + 2.0 * yk * T1[j] - odd1,
                           4.0 * mf * T1[j] + 2.0 * yk * T2[j] - odd2)
    Tinv = np.linalg.inv(T)
    kk = np.arange(n)
    C = np.cos(math.pi * np.outer(kk, kk) / (n - 1))
    C[:, 1:n - 1] *= 2.0
    C /= math.sqrt(2.0 * (n - 1))
    trunc = np.ones(n)
    trunc[p.n_cheb_ic_max:] = 0.0
    return dict(r_ic=r_ic, O_r_ic=O_r_ic, cheb=T, dcheb=T1, d2cheb=T2,
                cnorm=math.sqrt(2.0 / (n - 1)), D1=(Tinv @ T1).T, D2=(Tinv @ T2).T,
                CT=C * trunc[None, :])


def fft_radices(n: int) -> list:
    """Factor n into radices 4, 2, 3, 5 (4s first)."""
    out = []
    while n % 4 == 0:
        out.append(4); n //= 4
    while n % 2 == 0:
        out.append(2); n //= 2
    while n % 3 == 0:
        out.append(3); n //= 3
    while n % 5 == 0:
        out.append(5); n //= 5
    if n != 1:
        raise ValueError("n_phi/2 must factor into 2, 3, 5")
    return out


def sht_layout(p: Params, maps: dict) -> dict:
    """Per (m, parity) table sizes/offsets and the analysis tile list."""
    H = p.n_theta // 2
    Hp = -(-H // 32) * 32
    info = np.zeros((p.n_m, 2, 4), np.int32)
    syn_off = ana_off = 0
    for mi in range(p.n_m):
        m = mi * p.minc
        for par in range(2):
            K = (p.l_max - m - par) // 2 + 1 if p.l_max - m - par >= 0 else 0
            K4 = -(-K // 4) * 4
            K16 = -(-K // 16) * 16
            info[mi, par] = (K, syn_off, ana_off, K16)
            syn_off += K4 * Hp
            ana_off += K16 * Hp
    if syn_off >= 2 ** 31 or ana_off >= 2 ** 31:
        raise ValueError("Legendre tables exceed 32-bit offsets")
    out = dict(H=H, Hp=Hp, mp_info=info.reshape(-1), syn_size=syn_off, ana_size=ana_off)
    for waves in (6, 4):
        tl = analysis_tiles(p, info, 16 * waves)
        out[f"ana_tiles_{waves}"] = tl.reshape(-1)
        out[f"n_ana_tiles_{waves}"] = len(tl)
    return out


def analysis_tiles(p: Params, info: np.ndarray, tile: int) -> np.ndarray:
    """Analysis work list: one workgroup per (m, parity, first degree index k0)
    with `tile` degrees (16 per wave), as 8 equal-length lists (list x holds
    the orders m = x mod 8; the kernel gives list x to the workgroups of XCD
    x), padded with m = n_m."""
    per_xcd = [[] for _ in range(8)]
    for m in range(p.n_m):
        for par in range(2):
            for k0 in range(0, int(info[m, par, 0]), tile):
                per_xcd[m % 8].append((m, par, k0))
    tpx = max(len(t) for t in per_xcd)
    tiles = []
    for t in per_xcd:
        tiles += t + [(p.n_m, 0, 0)] * (tpx - len(t))
    return np.asarray(tiles, np.int32)
