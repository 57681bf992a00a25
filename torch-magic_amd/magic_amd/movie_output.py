"""Movie files in the reference's format (``movie_output.py:120-287``, Fortran
MagIC's out_movie_file): big-endian Fortran sequential records, float32 data.

A movie is a surface of one field: "Br CMB", "Vr EQ", "T EQ", ... (field Br,
Bt, Bp, Vr, Vt, Vp or T/S; surface CMB, r = r_cmb, or EQ, one theta row for
every radius).  The header (10 records) and the frame layout (a time record,
then one record per field) are the reference's, because the files are read
by MagIC's own tools.  The frame data are the grid values of the radial
loop's inverse transforms (``step_time.py:955-1059``): the velocity and field
in sin(theta)-weighted form and the entropy, computed here by this build's
SHT from the run's spectral fields; radial components are scaled by 1/r^2
and the others by 1/r as the reference does (``movie_output.py:195-287``).
The reference writes, at step n, the grids of its last radial loop, i.e. of
the state before step n (``main.py:852-868``); ``main.run`` captures the state
there the same way.
"""
from __future__ import annotations

import math
import struct

import numpy as np
import torch

# field name -> (field type code, file prefix) (movie_output.py:67-77)
FIELD_TYPES = {"BR": (1, "Br"), "BT": (2, "Bt"), "BP": (3, "Bp"), "VR": (4, "Vr"),
               "VT": (5, "Vt"), "VP": (6, "Vp"), "T": (7, "T"), "TEMPERATURE": (7, "T"),
               "S": (7, "T")}
# surface name -> (surface code, file suffix) (movie_output.py:80-84)
SURFACES = {"CMB": (1, "CMB"), "EQ": (2, "EQU"), "EQUAT": (2, "EQU")}


def _record(f, payload: bytes):
    """One Fortran sequential unformatted record with big-endian markers."""
    n = struct.pack(">i", len(payload))
    f.write(n + payload + n)


def _theta_sorted(p) -> np.ndarray:
    from .grid import gauss_legendre
    x, _ = gauss_legendre(p.n_theta)
    return np.arccos(x)   # ascending colatitude


def equator_row(n_theta: int) -> int:
    """Grid row of the equatorial movie.  The reference takes the sorted index
    of grid row n_theta/2 - 1 and uses it as a grid row (``movie_output.py:193``),
    so this is the row it reads; grid rows interleave (north j, south j)."""
    q = n_theta // 2 - 1
    return q // 2 if q % 2 == 0 else n_theta - 1 - q // 2


def parse_movie_string(s: str, p):
    """'Br CMB' -> (field type, surface code, surface constant, file stem)
    (reference ``parse_movie_string``, ``movie_output.py:87-117``)."""
    parts = s.strip().upper().split()
    if len(parts) < 2:
        raise ValueError(f"Invalid movie string: '{s}' (need FIELD SURFACE)")
    if parts[0] not in FIELD_TYPES:
        raise ValueError(f"Unknown movie field: '{parts[0]}'. Supported: {list(FIELD_TYPES)}")
    if parts[1] not in SURFACES:
        raise ValueError(f"Unknown movie surface: '{parts[1]}'. Supported: {list(SURFACES)}")
    ftype, prefix = FIELD_TYPES[parts[0]]
    nsurf, suffix = SURFACES[parts[1]]
    const = 1.0 if nsurf == 1 else (180.0 / math.pi) * float(_theta_sorted(p)[p.n_theta // 2 - 1])
    return ftype, nsurf, const, f"{prefix}_{suffix}_mov"


def write_movie_header(f, p, rad, n_surface: int, const: float, field_types,
                       runid="MagIC-PyTorch"):
    """The 10 header records (``write_movie_header``, ``movie_output.py:120-171``);
    p: the run's parameters, rad: its radial functions (``r``)."""
    n_ic = (p.n_r_max + 1) // 2
    _record(f, struct.pack(">i", 3))
    _record(f, struct.pack(">2i", n_surface, len(field_types)))
    _record(f, struct.pack(">d", float(const)))
    _record(f, struct.pack(f">{len(field_types)}i", *field_types))
    _record(f, runid.encode("ascii")[:64].ljust(64, b" "))
    _record(f, struct.pack(">7i", p.n_r_max + n_ic, p.n_r_max, n_ic, p.n_theta, p.n_phi, 0,
                           p.minc))
    _record(f, struct.pack(">8f", p.ra, p.ek, p.pr, p.raxi, p.sc, p.prmag, p.radratio, 1.0))
    r = torch.as_tensor(np.asarray(rad["r"]), dtype=torch.float64)
    r_cmb = float(r[0])
    # outer-core radii over r_cmb, then the reference's linear inner-core radii
    r_oc = r.to(torch.float32) / r_cmb
    r_ic = torch.linspace(float(r[-1]) / r_cmb, 0.0, n_ic, dtype=torch.float32)
    _record(f, torch.cat([r_oc, r_ic]).numpy().astype(">f4").tobytes())
    _record(f, np.asarray(_theta_sorted(p), np.float32).astype(">f4").tobytes())
    phi = torch.linspace(0, 2 * math.pi * (1 - 1.0 / p.n_phi), p.n_phi, dtype=torch.float32)
    _record(f, phi.numpy().astype(">f4").tobytes())


def write_movie_frame(f, time: float, frames):
    """A time record and one float32 record per field (``movie_output.py:174-189``)."""
    _record(f, struct.pack(">f", float(time)))
    for a in frames:
        _record(f, np.asarray(a, np.float32).astype(">f4").tobytes())


def grid_component(run, field_type: int, levels=None) -> torch.Tensor:
    """FP64 grid (levels, n_theta, n_phi) of one movie field of the run's current
    spectral state, in the grid's interleaved theta order: sin(theta)-weighted
    velocity / field components (torpol_to_spat of (l(l+1) w, dw, z) or
    (l(l+1) b, db, aj)) or the entropy (scal_to_spat of s)."""
    from .sht import SHT
    sht = SHT(run.res)
    f = run.f
    sel = slice(None) if levels is None else levels
    if field_type == 7:
        return sht.scal_to_spat(f["s"][:, sel].contiguous())
    dLh = run.res.d["dLh"].reshape(-1, 1).to(torch.complex128)
    if field_type <= 3:
        q, s, t = f["b"], f["db"], f["aj"]
    else:
        q, s, t = f["w"], f["dw"], f["z"]
    out = sht.torpol_to_spat((dLh * q[:, sel]).contiguous(), s[:, sel].contiguous(),
                             t[:, sel].contiguous())
    return out[(field_type - 1) % 3]


def movie_frame(run, field_type: int, n_surface: int) -> np.ndarray:
    """The float32 frame of one movie from the run's current state
    (``extract_surface_field`` / ``extract_equat_field``, ``movie_output.py:195-287``)."""
    p, rad = run.p, run.res.rad
    radial = field_type in (1, 4)
    if n_surface == 1:   # r = r_cmb, theta in sorted order
        g = grid_component(run, field_type, slice(0, 1))[0]
        scale = 1.0 if field_type == 7 else float(rad["or2"][0] if radial else rad["or1"][0])
        data = g if field_type == 7 else scale * g
        arr = data.to(torch.float32).cpu().numpy()
        H = p.n_theta // 2
        k = np.arange(p.n_theta)
        rows = np.where(k < H, 2 * k, 2 * (p.n_theta - 1 - k) + 1)   # sorted -> grid row
        return arr[rows]
    if n_surface == 2:   # one theta row, every radius; float32 product as the reference
        row = equator_row(p.n_theta)
        arr = grid_component(run, field_type)[:, row, :].to(torch.float32).cpu().numpy()
        if field_type == 7:
            return arr
        sc = np.asarray(rad["or2"] if radial else rad["or1"]).astype(np.float32)
        return arr * sc[:, None]
    raise ValueError(f"movie surface {n_surface} is not supported")


class MovieWriter:
    """The movie files of a run: ``MovieWriter(run, ["Br CMB", "Vr EQ"], dir, tag)``
    opens one file per movie (``<Field>_<CMB|EQU>_mov.<tag>``) and writes the
    headers; ``capture()`` takes the frames of the current state and
    ``write(time)`` appends them (``main.py:463-479,852-868``)."""

    def __init__(self, run, movies, output_dir: str, tag: str):
        import os
        self.run = run
        self.specs, self.files, self.pending = [], [], None
        for ms in movies:
            ftype, nsurf, const, stem = parse_movie_string(ms, run.p)
            fh = open(os.path.join(output_dir, f"{stem}.{tag}"), "wb")
            write_movie_header(fh, run.p, run.res.rad, nsurf, const, [ftype])
            print(f"  Movie: {ms} -> {fh.name}")
            self.specs.append((ftype, nsurf))
            self.files.append(fh)

    def capture(self):
        self.pending = [movie_frame(self.run, ft, ns) for ft, ns in self.specs]

    def write(self, time: float):
        if self.pending is None:
            self.capture()
        for fh, fr in zip(self.files, self.pending):
            write_movie_frame(fh, time, [fr])
            fh.flush()
        self.pending = None

    def close(self):
        for fh in self.files:
            fh.close()
        self.files = []
