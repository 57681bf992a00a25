"""Finite-difference radial scheme and its banded implicit operators (config 3).

Host set-up in FP64 numpy, built directly in LINPACK band storage for all
degrees at once; the per-degree factorisations and the solves of all
right-hand sides run on the GPU (``mh_band_factor`` / ``mh_band_solve``).

* grid: the reference's stretched FD grid (``finite_differences.py:62-127``):
  boundary layers of ``n_b = floor((n-1)/(2(1+fd_stretch)))`` points whose
  spacing grows geometrically from ``fd_ratio`` times the bulk spacing;
* stencils: Fornberg weights (``finite_differences.py:20-59``) for the bulk
  (fd_order) and boundary (fd_order_bound) rows of d/dr .. d4/dr4, with the
  row ranges of ``finite_differences.py:279-331``, kept as diagonals;
* background: polytropic anelastic state (``radial_functions.py:66-160``);
* operators of degree l (band half-widths as ``update_s.py:103-108``,
  ``update_z.py:121-127``, ``update_w_doublecurl.py:112-118``): entropy
  ``I - w opr kappa (D2 + (beta + dLtemp0 + 2/r + dLkappa) D1 - L/r^2)`` with
  Dirichlet rows (``update_s.py:146-163``); toroidal ``L/r^2 (I - w visc (D2 +
  (dLvisc - beta) D1 - (...)))`` with no-slip or stress-free rows
  (``update_z.py:160-190``); double-curl poloidal ``-L/r^2/rho (D2 - beta D1 -
  L/r^2) + w L/r^2/rho visc (D4 + c3 D3 + c2 D2 + c1 D1 + c0)`` with w = 0 and
  dw/dr = 0 (or stress-free) rows (``update_w_doublecurl.py:153-223``); row
  max-norm preconditioning, and for W then column max-norm.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib
from .device import C128, F64, I32, ptr, require_gpu, stream_handle


def fd_weights(x: np.ndarray, k: int) -> np.ndarray:
    """Fornberg weights of the k-th derivative at 0 from values at points x."""
    n = len(x)
    w = np.zeros((k + 1, n))
    w[0, 0] = 1.0
    c1, c4 = 1.0, x[0]
    for i in range(1, n):
        mn = min(i, k)
        c2, c5, c4 = 1.0, c4, x[i]
        ks = np.arange(mn, 0, -1)
        for j in range(i):
            c3 = x[i] - x[j]
            c2 *= c3
            if j == i - 1:
                w[ks, i] = c1 * (ks * w[ks - 1, i - 1] - c5 * w[ks, i - 1]) / c2
                w[0, i] = -c1 * c5 * w[0, i - 1] / c2
            w[ks, j] = (c4 * w[ks, j] - ks * w[ks - 1, j]) / c3
            w[0, j] = c4 * w[0, j] / c3
        c1 = c2
    return w[k]


def stretched_grid(n: int, ricb: float, rcmb: float, stretch: float, ratio: float) -> np.ndarray:
    """r[0] = rcmb > ... > r[n-1] = ricb with geometrically refined boundary layers."""
    r = np.empty(n)
    r[0] = rcmb
    if ratio == 1.0:
        for i in range(1, n):
            r[i] = r[i - 1] - 1.0 / (n - 1)
    else:
        nb = int((n - 1) / (2.0 * (1.0 + stretch)))
        q = math.exp(math.log(ratio) / nb)
        shell = abs(ricb) > 10.0 * np.finfo(np.float64).eps
        n_bulk = n - 1 - (2 if shell else 1) * nb
        qn = 1.0
        for _ in range(nb):
            qn *= q
        h = 1.0 / (n_bulk + (2.0 if shell else 1.0) * q * (1.0 - qn) / (1.0 - q))
        for _ in range(nb):
            h *= q                          # finest spacing at the outer wall
        steps = []
        for _ in range(nb):
            steps.append(h)
            h /= q
        steps += [h] * n_bulk
        if shell:
            for _ in range(nb):
                h *= q
                steps.append(h)
        for i in range(1, n):
            r[i] = r[i - 1] - steps[i - 1]
    r[n - 1] = ricb
    return r


def derivative_bands(r: np.ndarray, order: int, order_boundary: int, hw: int) -> dict:
    """{k: (n, 2hw+1) diagonals of d^k/dr^k}: B[i, hw+d] = D_k[i, i+d]."""
    n, h, ob = len(r), order // 2, order_boundary
    out = {}
    for k in (1, 2, 3, 4):
        half = h if k <= 2 else h + 1          # bulk half-width
        nbd = h if k <= 2 else h + 1           # boundary rows per side
        edge = ob + k                          # boundary stencil points
        B = np.zeros((n, 2 * hw + 1))
        for i in range(n):
            if i < nbd:
                cols = np.arange(edge)
            elif i >= n - nbd:
                cols = np.arange(n - edge, n)
            elif k >= 3 and (i < 1 + h or i >= n - h - 1):
                continue                       # rows outside the bulk range stay zero
            else:
                cols = np.arange(i - half, i + half + 1)
            d = cols - i
            if d.min() < -hw or d.max() > hw:
                raise ValueError("stencil wider than the work band")
            B[i, d + hw] = fd_weights(r[cols] - r[i], k)
        out[k] = B
    return out


def polytropic(r, rcmb, ricb, radratio, strat, polind, g0, g1, g2) -> dict:
    """Polytropic reference state and its log-derivatives."""
    or1 = 1.0 / r
    or2 = or1 * or1
    rgrav = g0 + g1 * r / rcmb + g2 * rcmb ** 2 * or2
    diss = ((math.exp(strat / polind) - 1.0) / (rcmb - ricb)
            / (g0 + 0.5 * g1 * (1.0 + radratio) + g2 / radratio))
    temp0 = (1.0 + diss * rcmb * (g0 + 0.5 * g1 - g2)
             - diss * (g0 * r + 0.5 * g1 * r * r / rcmb - g2 * rcmb ** 2 * or1))
    dgrav = g1 / rcmb - 2.0 * g2 * rcmb ** 2 * or2 * or1
    beta = -polind * diss * rgrav / temp0
    dbeta = -polind * diss * (dgrav * temp0 + diss * rgrav ** 2) / temp0 ** 2
    ddbeta = -polind * diss / temp0 ** 3 * (
        temp0 * (6.0 * g2 * rcmb ** 2 * or2 * or2 * temp0 + diss * rgrav * dgrav)
        + 2.0 * diss * rgrav * (diss * rgrav ** 2 + dgrav * temp0))
    n = len(r)
    z, o = np.zeros(n), np.ones(n)
    return dict(or1=or1, or2=or2, rho0=temp0 ** polind, orho1=temp0 ** -polind, beta=beta,
                dbeta=dbeta, ddbeta=ddbeta, dLtemp0=-diss * rgrav / temp0, visc=o, dLvisc=z,
                ddLvisc=z, kappa=o, dLkappa=z)


def _band_widths(order, order_boundary):
    if order <= 2 and order_boundary <= 2:
        return 1, (order + 2) // 2
    return max(order // 2, order_boundary), (max(order + 3, 2 * order_boundary + 3) - 1) // 2


class FDOperators:
    """Banded S, Z and W operators of every degree for one FD configuration.

    Built in LINPACK storage abd[kl+ku+i-j, j] = A[i, j] with 2kl+ku+1 rows
    (kl = ku), preconditioned; ``factor()`` runs the batched band LU on the GPU.
    """

    def __init__(self, n_r: int, l_max: int, wimp: float, *, radratio=0.35, fd_order=4,
                 fd_order_bound=2, fd_stretch=0.3, fd_ratio=0.1, strat=3.0, polind=2.0,
                 g0=0.0, g1=0.0, g2=1.0, opr=1.0, ktopv=2, kbotv=2):
        self.n, self.l_max = n_r, l_max
        rcmb = 1.0 / (1.0 - radratio)
        ricb = rcmb - 1.0
        self.r = r = stretched_grid(n_r, ricb, rcmb, fd_stretch, fd_ratio)
        self.kls, self.klw = _band_widths(fd_order, fd_order_bound)
        hw = 2 * fd_order_bound + 8               # work band, wider than any stencil
        D = derivative_bands(r, fd_order, fd_order_bound, hw)
        prof = polytropic(r, rcmb, ricb, radratio, strat, polind, g0, g1, g2) if strat > 0 else \
            dict(or1=1.0 / r, or2=1.0 / r ** 2, orho1=np.ones(n_r), beta=np.zeros(n_r),
                 dbeta=np.zeros(n_r), ddbeta=np.zeros(n_r), dLtemp0=np.zeros(n_r),
                 visc=np.ones(n_r), dLvisc=np.zeros(n_r), ddLvisc=np.zeros(n_r),
                 kappa=np.ones(n_r), dLkappa=np.zeros(n_r), rho0=np.ones(n_r))
        self.prof = prof
        I = np.zeros((n_r, 2 * hw + 1))
        I[:, hw] = 1.0
        ls = np.arange(l_max + 1, dtype=np.float64)
        dL = (ls * (ls + 1.0))[:, None, None]          # (L, 1, 1)
        col = lambda v: v[None, :, None]               # (1, n, 1)
        or1, or2, beta = prof["or1"], prof["or2"], prof["beta"]
        # entropy
        d1c = prof["beta"] + prof["dLtemp0"] + 2.0 * or1 + prof["dLkappa"]
        S = I[None] - wimp * opr * col(prof["kappa"]) * (
            D[2][None] + col(d1c) * D[1][None] - dL * col(or2) * I[None])
        S[:, 0], S[:, -1] = I[0], I[-1]
        self.S = self._finish(S, hw, self.kls, column=False)
        # toroidal (l = 0: identity, unused)
        vis, dLv, dbeta = prof["visc"], prof["dLvisc"], prof["dbeta"]
        c0z = dLv * beta + 2.0 * dLv * or1 + dbeta + 2.0 * beta * or1
        Z = dL * col(or2) * (I[None] - wimp * col(vis) * (
            D[2][None] + col(dLv - beta) * D[1][None] - (col(c0z) + dL * col(or2)) * I[None]))
        Z[:, 0] = (D[1][0] - (2.0 * or1[0] + beta[0]) * I[0]) if ktopv == 1 else I[0]
        Z[:, -1] = (D[1][-1] - (2.0 * or1[-1] + beta[-1]) * I[-1]) if kbotv == 1 else I[-1]
        Z[0] = I
        self.Z = self._finish(Z, hw, self.kls, column=False)
        # double-curl poloidal (l = 0: identity, unused)
        dbe, ddbe, ddv, orho = prof["dbeta"], prof["ddbeta"], prof["ddLvisc"], prof["orho1"]
        c3 = 2.0 * (dLv - beta)
        c2 = (ddv - 2.0 * dbe + dLv ** 2 + beta ** 2 - 3.0 * dLv * beta - 2.0 * or1 * (dLv + beta)
              )[None, :, None] - 2.0 * dL * col(or2)
        c1 = (-ddbe - dbe * (2.0 * dLv - beta + 2.0 * or1) - ddv * (beta + 2.0 * or1)
              + beta ** 2 * (dLv + 2.0 * or1) - beta * (dLv ** 2 - 2.0 * or2)
              - 2.0 * dLv * or1 * (dLv - or1))[None, :, None] \
            + dL * col(2.0 * (2.0 * or1 + beta - dLv) * or2)
        c0 = dL * col(or2) * (col(2.0 * dbe + ddv + dLv ** 2 - (2.0 / 3.0) * beta ** 2 + dLv * beta
                                  + 2.0 * or1 * (2.0 * dLv - beta - 3.0 * or1)) + dL * col(or2))
        old = -dL * col(or2 * orho) * (D[2][None] - col(beta) * D[1][None] - dL * col(or2) * I[None])
        imp = dL * col(orho * vis * or2) * (D[4][None] + col(c3) * D[3][None] + c2 * D[2][None]
                                            + c1 * D[1][None] + c0 * I[None])
        W = old + wimp * imp
        # boundary rows: row 0 / n-1 hold w = 0; row 1 / n-2 the velocity condition
        # at the wall (its stencil is the derivative row of the wall point,
        # shifted by one column offset)
        W[:, 0], W[:, -1] = I[0], I[-1]
        top = (D[2][0] - (2.0 * or1[0] + beta[0]) * D[1][0]) if ktopv == 1 else D[1][0]
        bot = (D[2][-1] - (2.0 * or1[-1] + beta[-1]) * D[1][-1]) if kbotv == 1 else D[1][-1]
        W[:, 1] = np.roll(top, -1)                  # D[0, j] sits at offset j - 1 in row 1
        W[:, -2] = np.roll(bot, 1)                  # D[n-1, j] sits at offset j - n + 2
        W[0] = I
        self.W = self._finish(W, hw, self.klw, column=True)
        self.dev = None

    @staticmethod
    def _finish(B, hw, k, column):
        """Precondition, check the band and convert (L, n, 2hw+1) diagonals to
        LINPACK storage (L, 3k+1, n); returns (abd, fac_row, fac_col)."""
        L, n, _ = B.shape
        fr = 1.0 / np.abs(B).max(axis=2)
        B = B * fr[:, :, None]
        fc = np.ones((L, n))
        i = np.arange(n)[:, None]
        d = np.arange(-hw, hw + 1)[None, :]
        j = i + d
        inside = (j >= 0) & (j < n)
        if column:
            cm = np.zeros((L, n))
            for dd in range(2 * hw + 1):
                rows = np.nonzero(inside[:, dd])[0]
                cols = rows + dd - hw                   # distinct for a fixed diagonal
                cm[:, cols] = np.maximum(cm[:, cols], np.abs(B[:, rows, dd]))
            fc = 1.0 / cm
            jj = np.clip(j, 0, n - 1)
            B = B * np.where(inside, fc[:, jj], 0.0)
        outside = (np.abs(d) > k) & inside
        if np.abs(B[:, outside]).max(initial=0.0) > 1e-14:
            raise ValueError(f"operator leaves the band kl=ku={k}")
        abd = np.zeros((L, 3 * k + 1, n))
        for dd in range(-k, k + 1):
            rows = np.arange(max(0, -dd), min(n, n - dd))
            abd[:, 2 * k - dd, rows + dd] = B[:, rows, dd + hw]
        return abd, fr, fc

    def factor(self, device="cuda"):
        """Batched LINPACK band LU of every degree and system on the GPU."""
        dev = require_gpu(device)
        self.dev = {}
        for name, (abd, fr, fc), k in (("S", self.S, self.kls), ("Z", self.Z, self.kls),
                                       ("W", self.W, self.klw)):
            a = torch.as_tensor(abd, dtype=F64).to(dev).contiguous()
            piv = torch.zeros(a.shape[0], self.n, dtype=I32, device=dev)
            info = torch.zeros(a.shape[0], dtype=I32, device=dev)
            _lib.call("mh_band_factor", ptr(a), ptr(piv), ptr(info), self.n, k, k, a.shape[0],
                      stream_handle())
            self.dev[name] = dict(abd=a, piv=piv, info=info, k=k,
                                  fr=torch.as_tensor(fr, dtype=F64).to(dev).contiguous(),
                                  fc=torch.as_tensor(fc, dtype=F64).to(dev).contiguous())
        return self

    def solve(self, name, rhs, l_index, out=None, rows=None):
        """Solve every row lm of rhs (L, n) complex with the degree l_index[lm]
        factors of system S, Z or W, all rows in one launch.  rows: the rows
        grouped by degree (computed from l_index when not given)."""
        s = self.dev[name]
        out = torch.empty_like(rhs) if out is None else out
        if rows is None:
            rows = torch.argsort(l_index.to(torch.int64), stable=True).to(I32)
        _lib.call("mh_band_solve", ptr(s["abd"]), ptr(s["piv"]), ptr(s["fr"]),
                  ptr(s["fc"]) if name == "W" else 0, ptr(l_index), ptr(rows), ptr(rhs),
                  ptr(out), rhs.shape[0], self.n, s["k"], s["k"], stream_handle())
        return out
