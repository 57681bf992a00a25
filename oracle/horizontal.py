"""Horizontal discretisation: lm maps, Gauss grid, Legendre tables (oracle).

Restates ``blocking.py:19-52`` (standard m-major lm order and l±1 maps),
``horizontal_data.py:16-52`` (Gauss-Legendre Newton iteration),
``horizontal_data.py:66-80`` (interleaved N/S theta layout),
``horizontal_data.py:88-137`` (clm, dLh, dTheta coupling arrays) and
``plms.py:21-186`` (orthonormal P_lm without Condon-Shortley phase and
sin(theta)·dP_lm/dtheta by the three-term recurrence).

Pinned by ``tests/test_oracle_grid.py``: nodes/weights against
``numpy.polynomial.legendre.leggauss``; P_lm against ``scipy.special`` values
(lpmv with the normalisation sqrt((2l+1)/4π (l-m)!/(l+m)!) and the
Condon-Shortley phase removed) and against the closed forms Y_0^0=1/sqrt(4π),
Y_1^0=sqrt(3/4π) cos θ; dP_lm against a centred finite difference in theta;
orthonormality sum_j w_j P_lm P_l'm = δ/(2π).
"""

from __future__ import annotations

import math

import numpy as np

from .params import Params


def lm_maps(p: Params):
    """m-major lm ordering (``blocking.py:19-52``).

    Returns lm2l, lm2m, lm2 (l,m)->lm, lm2lmS (l-1 or self), lm2lmA (l+1 or -1),
    and lm_start (first lm of each m).
    """
    L = p.lm_max
    lm2l = np.zeros(L, np.int64)
    lm2m = np.zeros(L, np.int64)
    lm2 = -np.ones((p.l_max + 1, p.l_max + 1), np.int64)
    lm_start = np.zeros(p.n_m_max + 1, np.int64)
    lm = 0
    for mc, m in enumerate(range(0, p.m_max + 1, p.minc)):
        lm_start[mc] = lm
        for l in range(m, p.l_max + 1):
            lm2l[lm], lm2m[lm], lm2[l, m] = l, m, lm
            lm += 1
    lm_start[p.n_m_max] = lm
    lm2lmS = np.arange(L)
    lm2lmA = -np.ones(L, np.int64)
    for lm in range(L):
        l, m = lm2l[lm], lm2m[lm]
        if l > 0 and l > m:
            lm2lmS[lm] = lm2[l - 1, m]
        if l < p.l_max:
            lm2lmA[lm] = lm2[l + 1, m]
    return dict(lm2l=lm2l, lm2m=lm2m, lm2=lm2, lm2lmS=lm2lmS, lm2lmA=lm2lmA,
                lm_start=lm_start)


def gauleg(n: int, x_min: float = -1.0, x_max: float = 1.0):
    """Gauss-Legendre colatitudes and weights (``horizontal_data.py:16-52``)."""
    eps = 10.0 * np.finfo(np.float64).eps
    theta = np.zeros(n)
    w = np.zeros(n)
    m = (n + 1) // 2
    xm = 0.5 * (x_max + x_min)
    xl = 0.5 * (x_max - x_min)
    for i in range(1, m + 1):
        z = math.cos(math.pi * (i - 0.25) / (n + 0.5))
        z1 = z + 10.0 * eps
        while abs(z - z1) > eps:
            p1, p2 = 1.0, 0.0
            for j in range(1, n + 1):
                p3 = p2
                p2 = p1
                p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j
            pp = n * (z * p1 - p2) / (z * z - 1.0)
            z1 = z
            z = z1 - p1 / pp
        theta[i - 1] = math.acos(xm + xl * z)
        theta[n - i] = math.acos(xm - xl * z)
        w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp)
        w[n - i] = w[i - 1]
    return theta, w


def grid_theta(p: Params):
    """Sorted and interleaved theta arrays (``horizontal_data.py:55-80``)."""
    theta_ord, gauss = gauleg(p.n_theta_max)
    nhs = p.n_theta_max // 2
    idx = np.zeros(p.n_theta_max, np.int64)
    idx[0::2] = np.arange(nhs)
    idx[1::2] = np.arange(p.n_theta_max - 1, nhs - 1, -1)
    theta_grid = theta_ord[idx]
    return dict(theta_ord=theta_ord, gauss=gauss, grid_idx=idx, theta_grid=theta_grid,
                sin_grid=np.sin(theta_grid), cos_grid=np.cos(theta_grid),
                O_sin_theta_E2_grid=1.0 / np.sin(theta_grid) ** 2)


def _clm(l, m):
    l = np.asarray(l, np.float64)
    m = np.asarray(m, np.float64)
    out = np.zeros(np.broadcast(l, m).shape)
    mask = l > 0
    lb = np.broadcast_to(l, out.shape)[mask]
    mb = np.broadcast_to(m, out.shape)[mask]
    out[mask] = np.sqrt((lb + mb) * (lb - mb) / ((2 * lb - 1) * (2 * lb + 1)))
    return out


def lm_arrays(p: Params, maps=None):
    """dLh, dPhi and dTheta coupling arrays (``horizontal_data.py:88-137``)."""
    maps = maps or lm_maps(p)
    l = maps["lm2l"].astype(np.float64)
    m = maps["lm2m"].astype(np.float64)
    c_l = _clm(l, m)
    c_lp1 = _clm(l + 1, m)
    out = dict(dLh=l * (l + 1), dPhi=1j * m)
    out["dTheta1S"] = (l + 1) * c_l
    out["dTheta1A"] = l * c_lp1
    out["dTheta2S"] = (l - 1) * c_l
    out["dTheta2A"] = (l + 2) * c_lp1
    out["dTheta3S"] = (l - 1) * (l + 1) * c_l
    out["dTheta3A"] = l * (l + 2) * c_lp1
    out["dTheta4S"] = out["dTheta1S"] * (l - 1) * l
    out["dTheta4A"] = out["dTheta1A"] * (l + 1) * (l + 2)
    return out


def plm_theta(theta: np.ndarray, l_max: int):
    """P_lm and sin(θ)·dP_lm/dθ at colatitudes ``theta`` (``plms.py:21-155``, norm=2).

    Vectorised over theta; loops over m and l exactly as the reference.
    Returns (plma, dplma), each (lm_max, len(theta)) in m-major lm order.
    """
    theta = np.atleast_1d(np.asarray(theta, np.float64))
    nt = theta.size
    lm_max = (l_max + 1) * (l_max + 2) // 2
    plma = np.zeros((lm_max, nt))
    dpl = np.zeros((lm_max, nt))
    cos_t = np.cos(theta)
    sin_t = np.sin(theta)
    dnorm = 1.0 / math.sqrt(4.0 * math.pi)
    pos = -1
    for m in range(0, l_max + 1):
        fac = 1.0
        for j in range(3, 2 * m + 2, 2):
            fac = fac * j / (j - 1)
        plm = np.full(nt, math.sqrt(fac))
        nz = sin_t != 0.0
        plm = np.where(nz, plm * sin_t ** m, 0.0 if m != 0 else plm)
        pos += 1
        plma[pos] = dnorm * plm
        plm1 = np.zeros(nt)
        for l in range(m + 1, l_max + 1):
            plm2 = plm1
            plm1 = plm
            plm = (cos_t * math.sqrt(float((2 * l - 1) * (2 * l + 1)) / float((l - m) * (l + m))) * plm1
                   - math.sqrt(float((2 * l + 1) * (l + m - 1) * (l - m - 1))
                               / float((2 * l - 3) * (l - m) * (l + m))) * plm2)
            pos += 1
            plma[pos] = dnorm * plm
        # extra P_{l_max+1}^m (plms.py:88-103), kept in dpl[pos] temporarily
        l = l_max + 1
        plm2 = plm1
        plm1 = plm
        if (l - m) * (l + m) > 0:
            plm = (cos_t * math.sqrt(float((2 * l - 1) * (2 * l + 1)) / float((l - m) * (l + m))) * plm1
                   - math.sqrt(float((2 * l + 1) * (l + m - 1) * (l - m - 1))
                               / float((2 * l - 3) * (l - m) * (l + m))) * plm2)
        else:
            plm = np.zeros(nt)
        dpl[pos] = dnorm * plm
    # sin(θ) dP/dθ (plms.py:105-153)
    pos = -1
    for m in range(0, l_max + 1):
        l = m
        pos += 1
        if m < l_max:
            dpl[pos] = l / math.sqrt(2 * l + 3) * plma[pos + 1]
        else:
            dpl[pos] = l / math.sqrt(2 * l + 3) * dpl[pos]
        for l in range(m + 1, l_max):
            pos += 1
            dpl[pos] = (l * math.sqrt(float((l + m + 1) * (l - m + 1)) / float((2 * l + 1) * (2 * l + 3)))
                        * plma[pos + 1]
                        - (l + 1) * math.sqrt(float((l + m) * (l - m)) / float((2 * l - 1) * (2 * l + 1)))
                        * plma[pos - 1])
        if m < l_max:
            l = l_max
            pos += 1
            dpl[pos] = (l * math.sqrt(float((l + m + 1) * (l - m + 1)) / float((2 * l + 1) * (2 * l + 3)))
                        * dpl[pos]
                        - (l + 1) * math.sqrt(float((l + m) * (l - m)) / float((2 * l - 1) * (2 * l + 1)))
                        * plma[pos - 1])
    return plma, dpl


def plm_tables(p: Params, grid=None):
    """Plm, dPlm, wPlm, wdPlm on the northern Gauss nodes (``plms.py:158-186``)."""
    grid = grid or grid_theta(p)
    nhs = p.n_theta_max // 2
    th = grid["theta_ord"][:nhs]
    w = grid["gauss"][:nhs]
    P, dP = plm_theta(th, p.l_max)
    if p.minc > 1:   # the rows of the orders m = 0, minc, ... in the run's lm order
        start = np.concatenate([[0], np.cumsum(p.l_max + 1 - np.arange(p.l_max + 1))])
        rows = np.concatenate([np.arange(start[m], start[m + 1])
                               for m in range(0, p.m_max + 1, p.minc)])
        P, dP = P[rows], dP[rows]
    return dict(Plm=P, dPlm=dP, wPlm=2.0 * math.pi * w * P, wdPlm=2.0 * math.pi * w * dP)
