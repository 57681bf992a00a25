"""Run parameters and derived truncation sizes (oracle restatement).

Follows ``src/magic_torch/params.py``: defaults at :57-190, the 2·3·5 prime
decomposition of the azimuthal grid at :12-24 and :69, ``lm_max`` at :76,
CNAB2's scheme-default CFL factors in ``time_scheme.py:13-19,33`` and the
derived physical factors of ``pre_calculations.py:21-32``.

Pinned by: ``tests/test_oracle_grid.py`` (sizes of Appendix-B style table:
l_max 32→n_phi 96, 64→192, 85→256, 127→384, 383→1152, lm_max closed form).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field


def prime_decomposition(n: int) -> int:
    """Smallest 2^i 3^j 5^k >= n (``params.py:12-24``)."""
    best, dist_min = n, 10 ** 9
    for i in range(13):
        for j in range(7):
            for k in range(7):
                val = 2 ** i * 3 ** j * 5 ** k
                dist = val - n
                if 0 <= dist < dist_min:
                    dist_min, best = dist, val
    return best


_DEFAULTS = dict(
    l_max=16, minc=1, n_r_max=None, n_cheb_max=None, n_cheb_ic_max=None,
    ra=1.0e5, ek=1.0e-3, pr=1.0, prmag=5.0, radratio=0.35, raxi=0.0, sc=10.0,
    mode=0, strat=0.0, polind=2.0, g0=0.0, g1=0.0, g2=0.0,
    ktops=1, kbots=1, ktopv=2, kbotv=2, kbotb=1,
    dtmax=1.0e-4, alpha=0.6, courfac=1e3, alffac=1e3, intfac=1e3,
    time_scheme="CNAB2", init_s1=404, amp_s1=0.1, init_v1=0,
    radial_scheme="chebyshev", fd_order=2, fd_order_bound=2,
    fd_stretch=0.3, fd_ratio=0.1,
    l_cour_alf_damp=True, sigma_ratio=0.0, nRotIC=0,
    l_correct_AMz=False, l_correct_AMe=False, omega_ic1=0.0,
)


@dataclass
class Params:
    """One run's configuration with the reference's derived quantities."""

    cfg: dict = field(default_factory=dict)

    def __post_init__(self):
        c = dict(_DEFAULTS)
        c.update(self.cfg)
        self.l_max = int(c["l_max"])
        self.minc = int(c["minc"])   # azimuthal symmetry: orders m = 0, minc, 2 minc, ...
        if self.minc < 1:
            raise ValueError(f"minc must be >= 1, got {self.minc}")
        self.m_max = self.l_max
        self.n_r_max = int(c["n_r_max"] if c["n_r_max"] is not None else 2 * self.l_max + 1)
        self.n_cheb_max = int(c["n_cheb_max"] if c["n_cheb_max"] is not None else self.n_r_max)
        # inner-core grid (params.py:65-66): (n_r_max + 1) / 2 points, even Chebyshev
        self.n_r_ic_max = (self.n_r_max + 1) // 2
        self.n_cheb_ic_max = int(c["n_cheb_ic_max"] if c["n_cheb_ic_max"] is not None
                                 else self.n_r_ic_max)
        nalias = 20
        # params.py:69-71: n_phi_max covers one 2 pi / minc sector
        self.n_phi_tot = prime_decomposition(2 * ((30 * self.l_max) // nalias))
        self.n_phi_max = self.n_phi_tot // self.minc
        self.n_theta_max = self.n_phi_tot // 2
        self.n_m_max = self.m_max // self.minc + 1
        self.lm_max = sum(self.l_max - m + 1 for m in range(0, self.m_max + 1, self.minc))
        for k in ("ra", "ek", "pr", "prmag", "radratio", "raxi", "sc", "strat", "polind",
                  "g0", "g1", "g2", "dtmax", "alpha", "courfac", "alffac", "intfac",
                  "amp_s1", "fd_stretch", "fd_ratio", "sigma_ratio", "omega_ic1"):
            setattr(self, k, float(c[k]))
        for k in ("mode", "ktops", "kbots", "ktopv", "kbotv", "kbotb", "init_s1",
                  "init_v1", "fd_order", "fd_order_bound", "nRotIC"):
            setattr(self, k, int(c[k]))
        self.time_scheme = str(c["time_scheme"])
        self.radial_scheme = str(c["radial_scheme"]).upper()
        self.l_cour_alf_damp = bool(c["l_cour_alf_damp"])
        self.l_correct_AMz = bool(c["l_correct_AMz"])
        self.l_correct_AMe = bool(c["l_correct_AMe"])
        self.l_mag = self.mode not in (1, 7)
        self.l_heat = self.mode not in (7,)
        self.l_anel = self.strat > 0.0
        self.l_finite_diff = self.radial_scheme == "FD"
        self.l_cond_ic = self.sigma_ratio > 0.0
        self.l_rot_ic = self.nRotIC != 0
        self.n_r_tot = self.n_r_max + self.n_r_ic_max if self.l_cond_ic else self.n_r_max
        self.O_sr = 1.0 / self.sigma_ratio if self.l_cond_ic else 0.0   # pre_calculations.py:52
        # pre_calculations.py:21-32 (tScale = lScale = 1)
        self.opr = 1.0 / self.pr
        self.opm = 1.0 / self.prmag
        self.CorFac = 1.0 / self.ek
        self.LFfac = 1.0 / (self.ek * self.prmag)
        self.BuoFac = self.ra / self.pr
        # time_scheme.py:13-19,33,184: scheme defaults of the CFL factors (CNAB2:
        # courfac 2.5, alffac 1.0, intfac 0.15; BPR353: 0.8, 0.35, 0.46), used
        # unless the configuration gives a value below 1e3
        if self.time_scheme not in ("CNAB2", "BPR353"):
            raise ValueError(f"Unknown time scheme: {self.time_scheme!r}")
        dc, da, di = (2.5, 1.0, 0.15) if self.time_scheme == "CNAB2" else (0.8, 0.35, 0.46)
        self.cnab2_courfac = dc if abs(self.courfac) >= 1e3 else self.courfac
        self.cnab2_alffac = da if abs(self.alffac) >= 1e3 else self.alffac
        self.cnab2_intfac = di if abs(self.intfac) >= 1e3 else self.intfac
        self.init_b1 = 3
        self.amp_b1 = 5.0

    @property
    def n_theta_half(self) -> int:
        return self.n_theta_max // 2


def sq4pi() -> float:
    return math.sqrt(4.0 * math.pi)
