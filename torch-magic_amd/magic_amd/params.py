"""Run configuration: the reference's YAML / ``configure`` keys and derived sizes.

Accepts every key of the reference's ``params.py:56-190`` (Appendix of the
reference config surface: l_max, minc, n_r_max, n_cheb_max, ra, ek, pr, prmag,
radratio, ..., radial_scheme, fd_order, ...) plus the run keys of
``main.py:299-308`` (n_steps, dt, log_every, ...).  Unknown keys raise
``KeyError`` so typos do not pass silently.  Values may come as strings (YAML
or ``MAGIC_*`` environment variables), as in the reference.

Derived sizes follow the reference: n_phi_tot is the smallest 2^a 3^b 5^c >=
2*floor(30 l_max / 20) (``params.py:12-24,69``), n_theta = n_phi_tot / 2, the
grid holds one 2 pi / minc sector of n_phi = n_phi_tot / minc longitudes,
lm_max = sum over the orders m = 0, minc, ... of (l_max - m + 1).
"""

from __future__ import annotations

import math
import os

PARAM_DEFAULTS = {
    "l_max": 16, "minc": 1, "n_r_max": None, "n_cheb_max": None, "n_cheb_ic_max": None,
    "ra": 1.0e5, "ek": 1.0e-3, "pr": 1.0, "prmag": 5.0, "radratio": 0.35, "raxi": 0.0,
    "sc": 10.0, "mode": 0, "strat": 0.0, "polind": 2.0, "g0": 0.0, "g1": 0.0, "g2": 0.0,
    "ktops": 1, "kbots": 1, "ktopxi": 1, "kbotxi": 1, "ktopv": 2, "kbotv": 2, "kbotb": 1,
    "dtmax": 1.0e-4, "alpha": 0.6, "courfac": 1e3, "alffac": 1e3, "intfac": 1e3,
    "radial_chunk_size": 0, "profile": False, "polar_opt": True, "use_shtns": True,
    "time_scheme": "CNAB2", "init_v1": 0, "init_s1": 404, "amp_s1": 0.1, "start_file": "",
    "radial_scheme": "chebyshev", "fd_order": 2, "fd_order_bound": 2, "fd_stretch": 0.3,
    "fd_ratio": 0.1, "l_correct_AMz": False, "l_correct_AMe": False, "l_cour_alf_damp": True,
    "sigma_ratio": 0.0, "nRotIC": 0, "n_log_step": 1, "l_power": False, "l_hel": False,
    "l_hemi": False, "alph1": 0.8, "alph2": 0.0, "stef": 0.0, "omega_ic1": 0.0,
    "device": "cuda",
    # this build's keys
    "precision": "fp64", "noise_amp": 0.0, "noise_seed": 0,
}

RUN_KEYS = {"n_steps": 100, "dt": None, "log_every": 1, "checkpoint_every": 0,
            "output_dir": ".", "restart": "", "fortran_restart": "", "n_graphs": 0,
            "movies": None, "n_movie_step": 0, "tag": "test"}

_INT = {"l_max", "minc", "n_r_max", "n_cheb_max", "n_cheb_ic_max", "mode", "ktops", "kbots",
        "ktopxi", "kbotxi", "ktopv", "kbotv", "kbotb", "radial_chunk_size", "init_v1",
        "init_s1", "fd_order", "fd_order_bound", "nRotIC", "n_log_step", "noise_seed"}
_BOOL = {"profile", "polar_opt", "use_shtns", "l_correct_AMz", "l_correct_AMe",
         "l_cour_alf_damp", "l_power", "l_hel", "l_hemi"}
_STR = {"time_scheme", "start_file", "radial_scheme", "device", "precision"}

# Arithmetic of the step (the one key added to the reference's): "fp64" (the
# reference's), or mixed precision with the Legendre transforms on FP32 MFMA
# ("mixed"), on BF16 MFMA with operands split into two bf16 planes, three
# products ("mixed_bf16"), or three planes, six products ("mixed_bf16x3").
PRECISIONS = ("fp64", "mixed", "mixed_bf16", "mixed_bf16x3")
LEGENDRE_PREC = {"fp64": 0, "mixed": 1, "mixed_bf16": 2, "mixed_bf16x3": 3}


def _as_bool(v) -> bool:
    return v if isinstance(v, bool) else str(v).lower() in ("true", "1", "yes")


def smallest_235(n: int) -> int:
    """Smallest 2^a 3^b 5^c >= n."""
    best = None
    p2 = 1
    while p2 < 4 * n + 8:
        p3 = p2
        while p3 < 4 * n + 8:
            p5 = p3
            while p5 < 4 * n + 8:
                if p5 >= n and (best is None or p5 < best):
                    best = p5
                p5 *= 5
            p3 *= 3
        p2 *= 2
    return best


class Params:
    """Resolved parameters of one dynamo run (immutable after construction)."""

    def __init__(self, cfg: dict | None = None, env: bool = True):
        cfg = dict(cfg or {})
        if "dt" in cfg and "dtmax" not in cfg and cfg["dt"] is not None:
            cfg["dtmax"] = cfg["dt"]          # run.py:24-26 alias
        vals = {}
        for k, d in PARAM_DEFAULTS.items():
            v = cfg.pop(k, None)
            if v is None and env:
                v = os.environ.get("MAGIC_" + k.upper())
            vals[k] = d if v is None else v
        self.run = dict(RUN_KEYS)
        for k in list(cfg):
            if k in RUN_KEYS:
                self.run[k] = cfg.pop(k)
        if cfg:
            raise KeyError(f"unknown configuration keys: {sorted(cfg)}")
        for k, v in vals.items():
            if v is None:
                setattr(self, k, None)
            elif k in _INT:
                setattr(self, k, int(v))
            elif k in _BOOL:
                setattr(self, k, _as_bool(v))
            elif k in _STR:
                setattr(self, k, str(v))
            else:
                setattr(self, k, float(v))
        self.radial_scheme = self.radial_scheme.upper()
        if self.minc < 1:
            raise ValueError(f"minc must be >= 1, got {self.minc}")
        if self.n_r_max is None:
            self.n_r_max = 2 * self.l_max + 1
        if self.n_cheb_max is None:
            self.n_cheb_max = self.n_r_max
        # inner core (params.py:65-66,162-173): (n_r_max + 1) / 2 even-Chebyshev
        # points; with a conducting inner core the B and J systems span both shells
        self.n_r_ic_max = (self.n_r_max + 1) // 2
        if self.n_cheb_ic_max is None:
            self.n_cheb_ic_max = self.n_r_ic_max
        # azimuthal symmetry minc (params.py:60-76): the orders m = 0, minc, ...
        # (n_m of them), the grid one 2 pi / minc sector of n_phi = n_phi_tot /
        # minc longitudes; n_theta = n_phi_tot / 2 whatever minc
        self.m_max = self.l_max
        self.n_m = self.l_max // self.minc + 1
        self.n_phi_tot = smallest_235(2 * ((30 * self.l_max) // 20))
        self.n_theta = self.n_phi_tot // 2
        if self.n_phi_tot % (2 * self.minc):
            raise ValueError(f"minc={self.minc} does not divide n_phi_tot={self.n_phi_tot} into "
                             "sectors of an even number of longitudes")
        self.n_phi = self.n_phi_tot // self.minc
        if self.n_theta % 2:
            raise ValueError(f"l_max={self.l_max} gives an odd number of latitudes "
                             f"({self.n_theta}); the N/S paired grid needs an even count")
        self.lm_max = sum(self.l_max - m + 1 for m in range(0, self.m_max + 1, self.minc))
        self.l_mag = self.mode not in (1, 7)
        self.l_heat = self.mode != 7
        self.l_anel = self.strat > 0.0
        self.l_finite_diff = self.radial_scheme == "FD"
        self.l_cond_ic = self.sigma_ratio > 0.0
        self.l_rot_ic = self.nRotIC != 0
        self.n_r_tot = self.n_r_max + self.n_r_ic_max if self.l_cond_ic else self.n_r_max
        self.O_sr = 1.0 / self.sigma_ratio if self.l_cond_ic else 0.0
        self.opr = 1.0 / self.pr
        self.opm = 1.0 / self.prmag
        self.CorFac = 1.0 / self.ek
        self.LFfac = 1.0 / (self.ek * self.prmag)
        self.BuoFac = self.ra / self.pr
        # the time scheme's CFL defaults unless the namelist value is below 1000
        # (time_scheme.py:13-19,33,184: CNAB2 2.5, 1.0, 0.15; BPR353 0.8, 0.35, 0.46)
        dc, da, di = (0.8, 0.35, 0.46) if self.time_scheme.upper() == "BPR353" else (2.5, 1.0, 0.15)
        self.cour_fac = dc if abs(self.courfac) >= 1e3 else self.courfac
        self.alf_fac = da if abs(self.alffac) >= 1e3 else self.alffac
        self.int_fac = di if abs(self.intfac) >= 1e3 else self.intfac
        self.amp_b1 = 5.0

    def check_step_supported(self):
        """The device step covers the Chebyshev path with CNAB2 or BPR353: the
        Boussinesq step with an insulating, conducting or rotating conducting
        inner core (nRotIC = 1: omega_ic integrated in time, no-slip ICB), and
        the anelastic step (strat > 0, polytropic; mixed precision in the
        Legendre transforms only) with an insulating inner core; angular-momentum corrections without a rotating core; the
        finite-difference radial scheme (double-curl poloidal equation),
        Boussinesq (any inner core) or anelastic (insulating inner core)."""
        bad = []
        if self.l_anel and self.l_cond_ic:
            bad.append("anelastic step with a conducting inner core")
        if self.l_anel and self.g0 == 0.0 and self.g1 == 0.0 and self.g2 == 0.0:
            bad.append("anelastic step without gravity (g0 = g1 = g2 = 0)")
        if self.nRotIC not in (0, 1):
            bad.append(f"nRotIC={self.nRotIC} (prescribed inner-core rotation)")
        elif self.nRotIC == 1 and not (self.l_cond_ic and self.l_mag and self.kbotv == 2):
            bad.append("rotating inner core without a conducting inner core, a magnetic "
                       "field and a no-slip ICB")
        if self.raxi != 0.0:
            bad.append("composition")
        if self.time_scheme.upper() not in ("CNAB2", "BPR353"):
            bad.append(f"time scheme {self.time_scheme}")
        if (self.l_correct_AMz or self.l_correct_AMe) and self.nRotIC != 0:
            bad.append("angular-momentum correction with a rotating inner core")
        if self.ktops != 1 or self.kbots != 1:
            bad.append("fixed-flux entropy boundaries")
        if self.precision not in PRECISIONS:
            bad.append(f"precision {self.precision!r} ({', '.join(PRECISIONS)})")
        if bad:
            raise NotImplementedError("device step does not support: " + ", ".join(bad))

    @property
    def sq4pi(self) -> float:
        return math.sqrt(4.0 * math.pi)

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k in PARAM_DEFAULTS}
        d.update(self.run)
        return d
