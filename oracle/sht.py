"""Spherical harmonic transforms (oracle restatement, FP64, numpy).

Restates the un-bucketed maths of ``sht.py``: ``scal_to_spat`` (:298-348),
``scal_to_SH`` (:351-405), ``torpol_to_spat`` (:408-497) and
``spat_to_sphertor`` (:575-667).  Per azimuthal order m the Legendre sums are
dense products with the northern-hemisphere tables; the southern hemisphere
uses the parity sign (-1)^(l-m) (``sht.py:111-118``); the azimuthal direction
is ``numpy.fft.irfft/rfft`` with norm="forward" (``sht.py:341,371``).
Grid layout is (batch, n_theta, n_phi) with interleaved N/S theta rows
(``horizontal_data.py:66-80``).

Pinned by ``tests/test_oracle_sht.py``: brute-force summation of
sum_lm P_lm(θ) S_lm e^{imφ} (with the real-field convention) at l_max<=16,
round trip analysis(synthesis(x)) = x to 1e-13, the vector synthesis against
a finite-difference gradient/curl of a scalar synthesis, and
spat_to_sphertor(torpol_to_spat(0,S,T)) = (S,T).
"""

from __future__ import annotations

import numpy as np

from .params import Params
from . import horizontal


class SHT:
    """Per-resolution SHT tables and the four transforms."""

    def __init__(self, p: Params, grid=None, tables=None, maps=None):
        self.p = p
        self.grid = grid or horizontal.grid_theta(p)
        self.maps = maps or horizontal.lm_maps(p)
        t = tables or horizontal.plm_tables(p, self.grid)
        self.nhs = p.n_theta_max // 2
        self.lm_start = self.maps["lm_start"]
        self.P, self.dP, self.wP, self.wdP = t["Plm"], t["dPlm"], t["wPlm"], t["wdPlm"]
        self.dLh = horizontal.lm_arrays(p, self.maps)["dLh"]
        self.ost2 = self.grid["O_sin_theta_E2_grid"]

    def _m_slices(self):
        p = self.p
        for mc in range(p.n_m_max):
            m = mc * p.minc
            s, e = self.lm_start[mc], self.lm_start[mc + 1]
            nlm = e - s
            sign = np.ones(nlm)
            sign[1::2] = -1.0           # (-1)^(l-m)
            yield m, slice(s, e), sign

    def _irfft(self, tmp):
        """tmp (n_theta, n_m, B) complex → (B, n_theta, n_phi) real, norm=forward."""
        p = self.p
        full = np.zeros((p.n_theta_max, p.n_phi_max // 2 + 1, tmp.shape[2]), np.complex128)
        full[:, :p.n_m_max, :] = tmp
        out = np.fft.irfft(full, n=p.n_phi_max, axis=1, norm="forward")
        return np.ascontiguousarray(out.transpose(2, 0, 1))

    def scal_to_spat(self, Slm: np.ndarray) -> np.ndarray:
        """(L,) or (L,B) complex → (n_theta,n_phi) or (B,n_theta,n_phi) (``sht.py:298``)."""
        batched = Slm.ndim == 2
        S = Slm if batched else Slm[:, None]
        B = S.shape[1]
        tmp = np.zeros((self.p.n_theta_max, self.p.n_m_max, B), np.complex128)
        for mc, (m, sl, sign) in enumerate(self._m_slices()):
            P = self.P[sl]                      # (nlm, H)
            tmp[0::2, mc] = P.T @ S[sl]
            tmp[1::2, mc] = (P * sign[:, None]).T @ S[sl]
        out = self._irfft(tmp)
        return out if batched else out[0]

    def scal_to_SH(self, sc: np.ndarray) -> np.ndarray:
        """(n_theta,n_phi) or (B,n_theta,n_phi) → (L,) or (L,B) (``sht.py:351``)."""
        batched = sc.ndim == 3
        f = sc if batched else sc[None]
        F = np.fft.rfft(f, axis=2, norm="forward")      # (B, nθ, nφ/2+1)
        out = np.zeros((self.p.lm_max, f.shape[0]), np.complex128)
        for mc, (m, sl, sign) in enumerate(self._m_slices()):
            fN = F[:, 0::2, mc].T                        # (H, B)
            fS = F[:, 1::2, mc].T
            wP = self.wP[sl]
            out[sl] = wP @ fN + (wP * sign[:, None]) @ fS
        return out if batched else out[:, 0]

    def torpol_to_spat(self, Q, S, T):
        """QST → (br, bt, bp) in sin(θ)-weighted form (``sht.py:408-497``)."""
        batched = Q.ndim == 2
        if not batched:
            Q, S, T = Q[:, None], S[:, None], T[:, None]
        B = Q.shape[1]
        bhG = S - 1j * T
        bhC = S + 1j * T
        bhG[0] = 0.0
        bhC[0] = 0.0
        nt = self.p.n_theta_max
        tr = np.zeros((nt, self.p.n_m_max, B), np.complex128)
        tt = np.zeros_like(tr)
        tp = np.zeros_like(tr)
        for mc, (m, sl, sign) in enumerate(self._m_slices()):
            P = self.P[sl]
            dP = self.dP[sl]
            PG = dP - m * P
            PC = dP + m * P
            s_es_neg = -sign                             # -1 for (l-m) even
            brN = P.T @ Q[sl]
            brS = (P * sign[:, None]).T @ Q[sl]
            N1 = PG.T @ bhG[sl]
            N2 = PC.T @ bhC[sl]
            S1 = (PC * s_es_neg[:, None]).T @ bhG[sl]
            S2 = (PG * s_es_neg[:, None]).T @ bhC[sl]
            tr[0::2, mc], tr[1::2, mc] = brN, brS
            tt[0::2, mc], tt[1::2, mc] = 0.5 * (N1 + N2), 0.5 * (S1 + S2)
            tp[0::2, mc], tp[1::2, mc] = -0.5j * (N1 - N2), -0.5j * (S1 - S2)
        out = tuple(self._irfft(t) for t in (tr, tt, tp))
        return out if batched else tuple(o[0] for o in out)

    def spat_to_sphertor(self, vt, vp):
        """(vθ, vφ) → (S, T) (``sht.py:575-667``)."""
        batched = vt.ndim == 3
        if not batched:
            vt, vp = vt[None], vp[None]
        ost2 = self.ost2[None, :, None]
        f2 = np.fft.rfft(vt, axis=2, norm="forward") * ost2
        f1 = np.fft.rfft(vp, axis=2, norm="forward") * ost2
        B = vt.shape[0]
        Sl = np.zeros((self.p.lm_max, B), np.complex128)
        Tl = np.zeros_like(Sl)
        for mc, (m, sl, sign) in enumerate(self._m_slices()):
            f1N, f1S = f1[:, 0::2, mc].T, f1[:, 1::2, mc].T
            f2N, f2S = f2[:, 0::2, mc].T, f2[:, 1::2, mc].T
            wP = self.wP[sl]
            wdP = self.wdP[sl]
            D = [wdP @ f2N, wdP @ f1N, wdP @ f2S, wdP @ f1S]
            C = [(-m * wP) @ (1j * f1N), (-m * wP) @ (1j * f2N),
                 (-m * wP) @ (1j * f1S), (-m * wP) @ (1j * f2S)]
            sg = sign[:, None]
            Sl[sl] = (C[0] + D[0]) + sg * (C[2] - D[2])
            Tl[sl] = (-D[1] + C[1]) + sg * (D[3] + C[3])
        inv = np.zeros(self.p.lm_max)
        inv[1:] = 1.0 / self.dLh[1:]
        Sl *= inv[:, None]
        Tl *= inv[:, None]
        return (Sl, Tl) if batched else (Sl[:, 0], Tl[:, 0])
