"""Device resources of one resolution and launchers for the SHT kernels.

A :class:`Resolution` uploads the host set-up of :mod:`magic_amd.grid`, builds
the Legendre tables on the GPU (``mh_plm_tables``) and the azimuthal FFT plan,
and offers thin launchers around the C ABI that take torch device tensors.
torch is used for device memory and the current stream only.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib, grid
from .params import Params

F64 = torch.float64
C128 = torch.complex128
I32 = torch.int32


def ptr(t) -> int:
    return 0 if t is None else t.data_ptr()


def stream_handle() -> C.c_void_p:
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def require_gpu(device) -> torch.device:
    device = torch.device(device)
    if device.type != "cuda" or not torch.cuda.is_available():
        raise _lib.MagicHipError("magic_amd runs on the GPU only (no CPU fallback); "
                                 "no HIP device is available")
    _lib.load()
    return device


class Resolution:
    """Per-(l_max, n_r_max, radratio, n_cheb_max) tables resident on the device."""

    def __init__(self, p: Params, device="cuda"):
        self.device = dev = require_gpu(device)
        self.p = p
        self.maps = grid.lm_maps(p)
        self.th = grid.theta_grid(p)
        self.rad = grid.radial(p)
        self.lmc = grid.lm_coefficients(p, self.maps)
        self.lay = grid.sht_layout(p, self.maps)
        self.L, self.N, self.M = p.lm_max, p.n_r_max, p.n_m
        self.T, self.P, self.H, self.Hp = p.n_theta, p.n_phi, self.lay["H"], self.lay["Hp"]

        def up(a, dt=F64):
            return torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(dev)

        self.d = {}
        self.d["mp_info"] = up(self.lay["mp_info"], I32)
        self.d["lm_start"] = up(self.maps["lm_start"], I32)
        self.d["l"] = up(self.maps["l"], I32)
        self.d["m"] = up(self.maps["m"], I32)
        self.d["l_rows"] = up(self.maps["l_rows"], I32)
        for w in (6, 4):
            self.d[f"ana_tiles_{w}"] = up(self.lay[f"ana_tiles_{w}"], I32)
        for k, v in self.lmc.items():
            self.d[k] = up(v)
        for k in ("or1", "or2", "or3", "or4", "r2", "rgrav", "q", "valri2", "valhi2"):
            self.d[k] = up(self.rad[k])
        self.d["ost2_grid"] = up(self.th["ost2_grid"])
        # Legendre tables
        self.synP = torch.zeros(self.lay["syn_size"], dtype=F64, device=dev)
        self.synD = torch.zeros_like(self.synP)
        self.anaW = torch.zeros(self.lay["ana_size"], dtype=F64, device=dev)
        self.anaWd = torch.zeros_like(self.anaW)
        cos_n, sin_n, w_n = up(self.th["cos_n"]), up(self.th["sin_n"]), up(self.th["w_n"])
        _lib.call("mh_plm_tables", p.l_max, p.n_m, p.minc, self.H, self.Hp, ptr(cos_n), ptr(sin_n),
                  ptr(w_n), ptr(self.d["mp_info"]), ptr(self.synP), ptr(self.synD),
                  ptr(self.anaW), ptr(self.anaWd), stream_handle())
        # FFT plan: complex length Mh = n_phi/2, twiddles exp(-2 pi i k/Mh), exp(-2 pi i k/n_phi)
        self.Mh = p.n_phi // 2
        self.radix = grid.fft_radices(self.Mh)
        k = np.arange(self.Mh)
        tw = np.exp(-2j * np.pi * k / self.Mh)
        tw2 = np.exp(-2j * np.pi * k / p.n_phi)
        self.d["tw"] = up(np.stack([tw.real, tw.imag], -1).reshape(-1))
        self.d["tw2"] = up(np.stack([tw2.real, tw2.imag], -1).reshape(-1))
        self.fft_rb = max(1, min(16, 2048 // self.Mh))
        self.d["comp_scale_v"] = up(np.array([0, 0, 0, 1, 1, 1, 1, 1, 1]), I32)
        self.d["comp_scale_2"] = up(np.array([1, 1]), I32)
        # the tables are shared by every run of this resolution, each on its own
        # stream: make them complete before any other stream can read them
        torch.cuda.current_stream(dev).synchronize()

    def f32_tables(self):
        """FP32 copies of the four Legendre tables, shared by every run of the
        resolution.  Built eagerly (a run or transform in mixed precision asks
        for them when it is created): never inside a recorded launch plan, where
        the conversion would be replayed with every step instead of run once."""
        if getattr(self, "_t32", None) is None:
            if getattr(_lib.recording(), "plan", False):
                raise RuntimeError("f32_tables: the FP32 tables must be built before "
                                   "a launch plan is recorded")
            t = {}
            for k, src in (("synP", self.synP), ("synD", self.synD), ("anaW", self.anaW),
                           ("anaWd", self.anaWd)):
                dst = torch.empty(src.numel(), dtype=torch.float32, device=self.device)
                _lib.call("mh_tables_to_f32", ptr(src), ptr(dst), src.numel(), stream_handle())
                t[k] = dst
            torch.cuda.current_stream(self.device).synchronize()  # shared across streams
            self._t32 = t
        return self._t32

    def bf16_tables(self, np_: int):
        """BF16 split copies (np_ = 2 or 3 planes) of the four Legendre tables for
        the split-precision transforms, shared by every run of the resolution and
        built eagerly, never inside a recorded launch plan (as f32_tables)."""
        if np_ not in (2, 3):
            raise ValueError(f"bf16_tables: 2 or 3 planes, not {np_}")
        cache = self.__dict__.setdefault("_t16", {})
        if np_ not in cache:
            if getattr(_lib.recording(), "plan", False):
                raise RuntimeError("bf16_tables: the BF16 tables must be built before "
                                   "a launch plan is recorded")
            info = self.lay["mp_info"].reshape(-1, 4)
            mp16 = np.zeros_like(info)
            so = ao = 0
            for i, (K, _s, _a, K16) in enumerate(info):
                K32 = -(-int(K) // 32) * 32
                mp16[i] = (K32, so, ao, 0)
                so += np_ * self.Hp * K32
                ao += np_ * int(K16) * self.Hp
            if so >= 2 ** 31 or ao >= 2 ** 31:
                raise ValueError("BF16 Legendre tables exceed 32-bit offsets")
            t = dict(mp16=torch.as_tensor(mp16.reshape(-1), dtype=I32).to(self.device))
            for k, n in (("synP", so), ("synD", so), ("anaW", ao), ("anaWd", ao)):
                t[k] = torch.zeros(max(n, 1), dtype=torch.int16, device=self.device)
            _lib.call("mh_tables_to_bf16", ptr(self.synP), ptr(self.synD), ptr(self.anaW),
                      ptr(self.anaWd), ptr(self.d["mp_info"]), ptr(t["mp16"]), len(info), self.H,
                      self.Hp, np_, ptr(t["synP"]), ptr(t["synD"]), ptr(t["anaW"]),
                      ptr(t["anaWd"]), stream_handle())
            torch.cuda.current_stream(self.device).synchronize()  # shared across streams
            cache[np_] = t
        return cache[np_]

    def _tables(self, prec):
        """(table pointers, mp16 pointer) of a transform precision: 0 FP64, 1 FP32,
        2 and 3 BF16 split with that many planes."""
        if prec >= 2:
            t = self.bf16_tables(prec)
            return t, ptr(t["mp16"])
        if prec == 1:
            return self.f32_tables(), 0
        return dict(synP=self.synP, synD=self.synD, anaW=self.anaW, anaWd=self.anaWd), 0

    # ------------------------------------------------------------ launchers
    def fstrides(self, Bout):
        """(component stride, m stride) in doubles of a Fourier buffer F[c][m][t][b]."""
        ms = 2 * self.T * Bout
        return self.M * ms, ms

    def synth(self, groups, B, ldx, F, Bout, vector: bool, or2=None, prec=0, waves=0, nct=0):
        """Legendre synthesis.  groups: list of [(x1, x2, mode)] * (3 if vector else 1).

        prec 0: FP64 MFMA; 1: FP32 tables and MFMA with FP32 accumulation; 2, 3:
        BF16 split tables and inputs of that many planes on the BF16 MFMA, FP32
        accumulation.  waves: waves per workgroup (0: the library default; 4 or
        6, 4 only for BF16); nct: column tiles per workgroup (0/1, or 2 for the
        FP32 and BF16 vector synthesis); same results."""
        a = _lib.SynArgs()
        for g, grp in enumerate(groups):
            for c, (x1, x2, mode) in enumerate(grp):
                a.in_[g][c].x1 = ptr(x1)
                a.in_[g][c].x2 = ptr(x2)
                a.in_[g][c].mode = mode
        a.ngroups, a.B, a.ldx, a.H, a.Hp = len(groups), B, ldx, self.H, self.Hp
        a.mp_info, a.lm_start = ptr(self.d["mp_info"]), ptr(self.d["lm_start"])
        tab, a.mp16 = self._tables(prec)
        a.synP, a.synD = ptr(tab["synP"]), ptr(tab["synD"])
        a.prec = prec
        a.fprec = 1 if F.dtype == torch.float32 else 0  # FP32 Fourier buffer (needs prec 1)
        a.or2 = ptr(or2 if or2 is not None else self.d["or2"])
        a.F = ptr(F)
        a.c_stride, a.m_stride = self.fstrides(Bout)
        a.Bout = Bout
        a.waves, a.nct = int(waves), int(nct)
        a.minc = self.p.minc
        _lib.call("mh_leg_synth", 1 if vector else 0, _lib.ref(a), self.M, stream_handle())

    def analysis(self, F, Bin, B, comps, outS, outT, ldo, vector: bool, prec=0, waves=0):
        """Legendre analysis of F[c][m][t][b] (comp stride = M*T*Bin complex).

        comps: scalar -> list of component indices; vector -> list of (c_vtheta, c_vphi).
        outS/outT: output tensors (or pointers as ints) per component / pair.
        waves: waves per workgroup (0: the library default, 6; or 4), same results.
        """
        a = _lib.AnaArgs()
        a.F = ptr(F)
        a.fprec = 1 if F.dtype == torch.float32 else 0  # FP32 Fourier buffer (needs prec 1)
        a.c_stride, a.m_stride = self.fstrides(Bin)
        a.Bin, a.B, a.ncomp = Bin, B, len(comps)
        for i, c in enumerate(comps):
            if vector:
                a.cf2[i], a.cf1[i] = c
            else:
                a.cf1[i] = c
            a.outS[i] = outS[i] if isinstance(outS[i], int) else ptr(outS[i])
            if vector and outT[i] is not None:
                a.outT[i] = outT[i] if isinstance(outT[i], int) else ptr(outT[i])
        a.ldo, a.H, a.Hp = ldo, self.H, self.Hp
        a.mp_info, a.lm_start = ptr(self.d["mp_info"]), ptr(self.d["lm_start"])
        tab, a.mp16 = self._tables(prec)
        a.anaW, a.anaWd = ptr(tab["anaW"]), ptr(tab["anaWd"])
        a.prec = prec
        w = waves or (4 if prec >= 2 else 6)
        a.tiles, a.ntiles = ptr(self.d[f"ana_tiles_{w}"]), self.lay[f"n_ana_tiles_{w}"]
        a.waves = w
        a.n_m, a.minc = self.M, self.p.minc
        _lib.call("mh_leg_analysis", 1 if vector else 0, _lib.ref(a), stream_handle())

    def fft(self, forward: bool, F, FB, G, ncomp, B, comp_scale=None):
        """Azimuthal FFT between F[c][m][t][b] (FB columns) and G[c][b][t][phi]."""
        a = _lib.FftArgs()
        a.Mh, a.nstage = self.Mh, len(self.radix)
        for i, r in enumerate(self.radix):
            a.radix[i] = r
        a.n_m, a.T, a.B, a.ncomp, a.RB = self.M, self.T, B, ncomp, self.fft_rb
        a.tw, a.tw2 = ptr(self.d["tw"]), ptr(self.d["tw2"])
        a.F = ptr(F)
        a.Fm_stride = 2 * self.T * FB
        a.Fc_stride = self.M * a.Fm_stride
        a.FB = FB
        a.G = ptr(G)
        a.Gc_stride = B * self.T * self.P
        a.comp_scale = ptr(comp_scale)
        a.ost2 = ptr(self.d["ost2_grid"])
        _lib.call("mh_fft", 1 if forward else 0, _lib.ref(a), stream_handle())

    def fourier_buffer(self, ncomp, B, dtype=F64):
        """F[c][m][t][b] complex as interleaved (re, im) of `dtype` (float64, or
        float32 for the mixed-precision radial loop)."""
        return torch.empty(ncomp * self.M * self.T * B * 2, dtype=dtype, device=self.device)

    def grid_step(self, Fs, Fv, Fo, N, cfl, p, generic=False, rb=0):
        """Fused inverse FFT + nonlinear products + Courant maxima + forward FFT.

        generic: use the kernel with a run-time FFT plan even where the library
        has one compiled for this n_phi; rb: radial levels per workgroup (0: the
        most the plan fits).  Every choice gives the same results.
        """
        a = _lib.GridArgs()
        a.Fs, a.Fv, a.Fo = ptr(Fs), ptr(Fv), ptr(Fo)
        if not (Fs.dtype == Fv.dtype == Fo.dtype):
            raise ValueError("grid_step: Fourier buffers of mixed dtypes")
        a.fprec = 1 if Fs.dtype == torch.float32 else 0
        a.N, a.T, a.M, a.Mh, a.generic = N, self.T, self.M, self.Mh, int(generic)
        a.tw, a.tw2 = ptr(self.d["tw"]), ptr(self.d["tw2"])
        d = self.d
        a.or2, a.or4, a.ost2 = ptr(d["or2"]), ptr(d["or4"]), ptr(d["ost2_grid"])
        a.valri2, a.valhi2 = ptr(d["valri2"]), ptr(d["valhi2"])
        a.LFfac, a.cf2, a.af2 = p.LFfac, p.cour_fac ** 2, p.alf_fac ** 2
        a.l_mag, a.alf_damp = int(p.l_mag), int(p.l_cour_alf_damp)
        a.cfl = ptr(cfl)
        a.rb = int(rb)
        _lib.call("mh_grid_step", _lib.ref(a), stream_handle())

    def grid_plan(self):
        """(register radix, threads per level, most levels per workgroup) of the
        grid step's FFT plan for this n_phi."""
        out = (C.c_int * 3)()
        if _lib.load().mh_grid_plan(self.Mh, C.cast(out, C.c_void_p)) != 0:  # host query, never recorded
            raise _lib.MagicHipError(f"mh_grid_plan: no FFT plan for n_phi = {2 * self.Mh}")
        return tuple(out)

    def grid_buffer(self, ncomp, B):
        return torch.empty(ncomp * B * self.T * self.P, dtype=F64, device=self.device)
