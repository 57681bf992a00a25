"""G_1.TAG graph files: grid-space fields of a run in MagIC's binary layout
(reference ``graph_output.py:53-242``, MagIC's ``out_graph_file.f90``).

The file is an exchange format, so its layout is MagIC's: big-endian float32
and int32 with no record markers; a header (version 14, a 64-character run
id, the time, 9 physical parameters, 5 grid integers, 6 flags, the
colatitudes in geographic order, the radii, and for a magnetic run the
inner-core radii); then per outer-core level u_r, u_theta, u_phi (zero on the
two rigid walls, as the reference writes them), s, p, B_r, B_theta, B_phi as
(n_theta, n_phi) column-major slabs in geographic theta order; then for a
magnetic run B_r, B_theta, B_phi of the potential field inside an insulating
inner core, level by level.

The grid values come from this library's synthesis kernels (all radial
levels in one launch per transform); the host converts them to float32 and
writes them.  This is a diagnostic file, off the step path.
"""

from __future__ import annotations

import math
import struct

import numpy as np
import torch

from .sht import SHT


def _theta_geo(p):
    from .grid import gauss_legendre
    x, _ = gauss_legendre(p.n_theta)
    return np.arccos(x)                       # ascending colatitude (north first)


def _cal2ord(T):
    """geo index g -> interleaved grid row: rows are (north j, south j) pairs."""
    H = T // 2
    g = np.arange(T)
    return np.where(g < H, 2 * g, 2 * (T - 1 - g) + 1)


def ic_radii(p, r_icb):
    """Inner-core radii (n_r_ic_max points of the even Chebyshev grid on
    [-r_icb, r_icb], the centre last) and their inverses (0 at the centre)
    (reference ``radial_functions.py:309-322``)."""
    n = (p.n_r_max + 1) // 2
    ntot = 2 * n - 1
    r = np.array([r_icb * math.cos(math.pi * k / (ntot - 1)) for k in range(n)])
    r[n - 1] = 0.0
    o = np.zeros(n)
    o[:n - 1] = 1.0 / r[:n - 1]
    return r, o


def write_graph_file(path, time: float, run, fields=None):
    """Write the graph file of `run` (its current fields, or `fields`: a dict
    of w, dw, z, s, p, b, db, aj arrays) at `time` to `path`.  The inner-core
    slabs hold the potential field continued from the outer b at the ICB, also
    for a conducting inner core, as the reference writes them
    (``graph_output.py:174-242``)."""
    p, res = run.p, run.res
    f = dict(run.f)
    if fields:
        f.update({k: torch.as_tensor(v).to(run.device, torch.complex128)
                  for k, v in fields.items()})
    T, P, N = p.n_theta, p.n_phi, p.n_r_max
    rad = res.rad
    n_ic = (N + 1) // 2
    out = bytearray()
    be_i = lambda v: out.extend(struct.pack(">i", int(v)))
    be_f = lambda v: out.extend(struct.pack(">f", float(v)))
    be_arr = lambda a: out.extend(np.asarray(a, dtype=">f4").tobytes())
    be_i(14)
    out.extend(b"Benchmark 1".ljust(64))
    be_f(time)
    for v in (p.ra, p.pr, p.raxi, p.sc, p.ek, 0.0, p.prmag, p.radratio, p.sigma_ratio):
        be_f(v)
    for v in (N, T, p.n_phi_tot, p.minc, n_ic):   # the full circle's count in the header
        be_i(v)
    for v in (p.l_heat, False, False, p.l_mag, True, p.l_cond_ic):
        be_i(1 if v else 0)
    be_arr(_theta_geo(p))
    be_arr(rad["r"])
    r_ic, o_ic = ic_radii(p, rad["r_icb"])
    if p.l_mag and n_ic > 1:
        be_arr(r_ic)

    sht = SHT(res)
    dLh = torch.as_tensor(res.lmc["dLh"], dtype=torch.complex128).to(run.device)[:, None]
    vr, vt, vp = sht.torpol_to_spat(dLh * f["w"], f["dw"], f["z"])
    s_g = sht.scal_to_spat(f["s"]) if p.l_heat else None
    p_g = sht.scal_to_spat(f["p"])
    if p.l_mag:
        br, bt, bp = sht.torpol_to_spat(dLh * f["b"], f["db"], f["aj"])
    c2o = torch.as_tensor(_cal2ord(T)).to(run.device)
    osn = torch.as_tensor(1.0 / res.th["sin_grid"]).to(run.device)[:, None]
    or1 = torch.as_tensor(rad["or1"]).to(run.device)
    or2 = torch.as_tensor(rad["or2"]).to(run.device)

    def slab(g):   # (T, P) interleaved rows -> geographic rows, column-major f32 bytes
        a = g[c2o].to(torch.float32).cpu().numpy()
        return a.astype(">f4").tobytes(order="F")

    zero = slab(torch.zeros(T, P, dtype=torch.float64, device=run.device))
    # anelastic runs write the velocity, rho0 u / rho0 (graph_output.py:144-147)
    orh = run.prof["orho1"] if getattr(run, "anel", False) else None
    for n in range(N):
        if n in (0, N - 1):
            out.extend(zero * 3)
        else:
            fr, ft = (or2[n], or1[n]) if orh is None else (or2[n] * orh[n], or1[n] * orh[n])
            out.extend(slab(fr * vr[n]))
            out.extend(slab(ft * osn * vt[n]))
            out.extend(slab(ft * osn * vp[n]))
        if p.l_heat:
            out.extend(slab(s_g[n]))
        out.extend(slab(p_g[n]))
        if p.l_mag:
            out.extend(slab(or2[n] * br[n]))
            out.extend(slab(or1[n] * osn * bt[n]))
            out.extend(slab(or1[n] * osn * bp[n]))
    if p.l_mag and n_ic > 1:
        # potential field below the inner-core boundary (graph_output.py:180-242):
        # Q = (r/r_icb)^(l+1) l(l+1) b_icb, S = (r/r_icb)^l (l+1)/r_icb b_icb, T = 0
        r_icb = rad["r_icb"]
        l = torch.as_tensor(res.maps["l"], dtype=torch.float64).to(run.device)[:, None]
        ratio = torch.as_tensor(r_ic / r_icb).to(run.device)[None, :]
        pw = torch.where(ratio > 0, ratio ** l, torch.zeros_like(ratio * l))
        b_icb = f["b"][:, N - 1:N]
        Q = (pw * ratio * l * (l + 1.0)) * b_icb
        S = (pw / r_icb * (l + 1.0)) * b_icb
        Bi = sht.torpol_to_spat(Q, S, torch.zeros_like(Q))
        o1 = torch.as_tensor(o_ic).to(run.device)
        for n in range(n_ic):
            out.extend(slab(o1[n] * o1[n] * Bi[0][n]))
            out.extend(slab(o1[n] * osn * Bi[1][n]))
            out.extend(slab(o1[n] * osn * Bi[2][n]))
    with open(path, "wb") as fh:
        fh.write(bytes(out))
