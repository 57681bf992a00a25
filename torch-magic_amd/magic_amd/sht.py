"""Spherical harmonic transforms with the reference's signatures, on the GPU.

``scal_to_spat`` / ``scal_to_SH`` / ``torpol_to_spat`` / ``spat_to_sphertor``
keep the shapes, layouts and conventions of the reference ``sht.py:298,351,
408,575`` (spectral (lm_max,) or (lm_max, B) complex128 in m-major order;
grid (n_theta, n_phi) or (B, n_theta, n_phi) float64 with interleaved N/S
theta rows; azimuthal FFTs with norm="forward"; vector fields in the
sin(theta)-weighted form), but run on the hand-written kernels of
libmagic_hip: FP64 MFMA Legendre contractions and LDS Stockham FFTs.

The module-level functions use the resolution of the configured run
(:func:`magic_amd.config.configure`); :class:`SHT` binds any resolution.

``lcut`` is accepted and has no effect, exactly as in the reference: its four
transforms set ``lcut = l_max`` when it is None and never read it again
(``sht.py:308-309,361-362,420-421,586-587``), so every degree up to l_max
always takes part; a drop-in keeps that result.
"""

from __future__ import annotations

import torch

from .device import C128, F64, Resolution


def _spec(x, L):
    x = torch.as_tensor(x)
    if x.dtype != C128:
        x = x.to(C128)
    batched = x.dim() == 2
    x = x if batched else x.unsqueeze(1)
    if x.shape[0] != L:
        raise ValueError(f"expected {L} spectral rows, got {x.shape[0]}")
    return x.contiguous(), batched


def _grid(x, T, P):
    x = torch.as_tensor(x)
    if x.dtype != F64:
        x = x.to(F64)
    batched = x.dim() == 3
    x = x if batched else x.unsqueeze(0)
    if tuple(x.shape[1:]) != (T, P):
        raise ValueError(f"expected grid ({T}, {P}), got {tuple(x.shape[1:])}")
    return x.contiguous(), batched


class SHT:
    """Transforms of one resolution.  precision "fp64" (the reference's),
    "mixed" (Legendre contractions on FP32 MFMA with FP32 tables), "mixed_bf16"
    or "mixed_bf16x3" (on BF16 MFMA with tables and inputs split into two or
    three bf16 planes); FP32 accumulation and FP64 FFTs in every mixed form;
    inputs and outputs keep the reference's dtypes."""

    def __init__(self, res: Resolution, precision: str = "fp64"):
        from .params import LEGENDRE_PREC, PRECISIONS
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {PRECISIONS}, not {precision!r}")
        self.res = res
        self.prec = LEGENDRE_PREC[precision]
        if self.prec == 1:
            res.f32_tables()
        elif self.prec >= 2:
            res.bf16_tables(self.prec)

    def _on(self, x):
        return x.to(self.res.device) if x.device != self.res.device else x

    def scal_to_spat(self, Slm, lcut=None):
        r = self.res
        S, batched = _spec(Slm, r.L)
        S = self._on(S)
        B = S.shape[1]
        F = r.fourier_buffer(1, B)
        r.synth([[(S, None, 0)]], B, B, F, B, vector=False, prec=self.prec)
        G = torch.empty(B, r.T, r.P, dtype=F64, device=r.device)
        r.fft(False, F, B, G, 1, B)
        return G if batched else G[0]

    def scal_to_SH(self, sc, lcut=None):
        r = self.res
        g, batched = _grid(sc, r.T, r.P)
        g = self._on(g)
        B = g.shape[0]
        F = r.fourier_buffer(1, B)
        r.fft(True, F, B, g, 1, B)
        out = torch.zeros(r.L, B, dtype=C128, device=r.device)
        r.analysis(F, B, B, [0], [out], [None], B, vector=False, prec=self.prec)
        return out if batched else out[:, 0]

    def torpol_to_spat(self, Qlm, Slm, Tlm, lcut=None):
        r = self.res
        Q, batched = _spec(Qlm, r.L)
        S, _ = _spec(Slm, r.L)
        T, _ = _spec(Tlm, r.L)
        Q, S, T = self._on(Q), self._on(S), self._on(T)
        B = Q.shape[1]
        F = r.fourier_buffer(3, B)
        r.synth([[(Q, None, 0), (S, None, 0), (T, None, 0)]], B, B, F, B, vector=True,
                prec=self.prec)
        G = torch.empty(3, B, r.T, r.P, dtype=F64, device=r.device)
        r.fft(False, F, B, G, 3, B)
        out = (G[0], G[1], G[2])
        return out if batched else tuple(o[0] for o in out)

    def spat_to_sphertor(self, vt, vp, lcut=None):
        r = self.res
        t, batched = _grid(vt, r.T, r.P)
        p, _ = _grid(vp, r.T, r.P)
        B = t.shape[0]
        G = torch.empty(2, B, r.T, r.P, dtype=F64, device=r.device)
        G[0].copy_(self._on(t))
        G[1].copy_(self._on(p))
        F = r.fourier_buffer(2, B)
        r.fft(True, F, B, G, 2, B, comp_scale=r.d["comp_scale_2"])
        S = torch.zeros(r.L, B, dtype=C128, device=r.device)
        T = torch.zeros(r.L, B, dtype=C128, device=r.device)
        r.analysis(F, B, B, [(0, 1)], [S], [T], B, vector=True, prec=self.prec)
        return (S, T) if batched else (S[:, 0], T[:, 0])


_default = None


def _sht() -> SHT:
    global _default
    if _default is None:
        from . import config
        _default = SHT(config.default_resolution())
    return _default


def scal_to_spat(Slm, lcut=None):
    """(lm_max,) or (lm_max, B) complex → grid (reference ``sht.py:298``)."""
    return _sht().scal_to_spat(Slm, lcut)


def scal_to_SH(sc, lcut=None):
    """Grid → (lm_max,) or (lm_max, B) complex (reference ``sht.py:351``)."""
    return _sht().scal_to_SH(sc, lcut)


def torpol_to_spat(Qlm, Slm, Tlm, lcut=None):
    """QST → (br, bt, bp) sin(theta)-weighted (reference ``sht.py:408``)."""
    return _sht().torpol_to_spat(Qlm, Slm, Tlm, lcut)


def spat_to_sphertor(vt, vp, lcut=None):
    """(vtheta, vphi) → (S, T) (reference ``sht.py:575``)."""
    return _sht().spat_to_sphertor(vt, vp, lcut)


def spat_to_SH(sc, lcut=None):
    return scal_to_SH(sc, lcut)
