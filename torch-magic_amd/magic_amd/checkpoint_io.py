"""Fortran MagIC stream checkpoints (versions 2-5) with the reference's reader
and writer behaviour (``checkpoint_io.py:100-309`` reader, ``:336-490`` writer).

The layout is MagIC's binary exchange format, so it is reproduced field for
field: version (int32), time, the time-scheme family (10 characters), nexp /
nimp / nold, the dt array, the step number, 9 physical parameters, 6 grid
integers, (v4+) l_max / m_min / m_max, the radial scheme (72 characters, two
ints, two doubles), the radial grid, (MULTISTEP) the rotation-rate history
scalars, 12 rotation doubles, the logical flags (4-byte), then each field as
a column-major complex (lm_max, n_r) array followed by its multistep history
arrays (explicit slots 1..nexp-1, implicit 1..nimp-1, old 1..nold-1).
Endianness is detected from the version word on reading; the writer emits
big-endian by default, as the reference does.

The device state goes through host memory once per file; nothing here is on
the step path.
"""

from __future__ import annotations

import math
import struct
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch


@dataclass
class CheckpointData:
    """Everything a MagIC checkpoint holds (fields in m-major lm order)."""
    version: int = 0
    time: float = 0.0
    family: str = ""
    nexp: int = 0
    nimp: int = 0
    nold: int = 0
    dt: np.ndarray = field(default_factory=lambda: np.zeros(0))
    n_time_step: int = 0
    ra: float = 0.0
    pr: float = 0.0
    raxi: float = 0.0
    sc: float = 0.0
    prmag: float = 0.0
    ek: float = 0.0
    stef: float = 0.0
    radratio: float = 0.0
    sigma_ratio: float = 0.0
    n_r_max: int = 0
    n_theta_max: int = 0
    n_phi_tot: int = 0
    minc: int = 0
    nalias: int = 0
    n_r_ic_max: int = 0
    l_max: int = 0
    m_min: int = 0
    m_max: int = 0
    n_cheb_max: int = 0
    r: np.ndarray = field(default_factory=lambda: np.zeros(0))
    omega_ic1: float = 0.0
    omega_ma1: float = 0.0
    l_heat: bool = False
    l_chemical_conv: bool = False
    l_phase_field: bool = False
    l_mag: bool = False
    l_press_store: bool = False
    l_cond_ic: bool = False
    w: Optional[np.ndarray] = None
    z: Optional[np.ndarray] = None
    p: Optional[np.ndarray] = None
    s: Optional[np.ndarray] = None
    xi: Optional[np.ndarray] = None
    phi: Optional[np.ndarray] = None
    b: Optional[np.ndarray] = None
    aj: Optional[np.ndarray] = None
    b_ic: Optional[np.ndarray] = None
    aj_ic: Optional[np.ndarray] = None
    derivatives: dict = field(default_factory=dict)
    domega_ic_dt: Optional[dict] = None
    domega_ma_dt: Optional[dict] = None
    endian: str = "<"


class _Reader:
    def __init__(self, buf: bytes, endian: str):
        self.buf, self.pos, self.e = buf, 0, endian

    def take(self, n: int) -> bytes:
        if self.pos + n > len(self.buf):
            raise ValueError(f"checkpoint ends at byte {len(self.buf)}, needed {self.pos + n}")
        out = self.buf[self.pos:self.pos + n]
        self.pos += n
        return out

    def unpack(self, fmt: str):
        s = struct.Struct(self.e + fmt)
        return s.unpack(self.take(s.size))

    def array(self, dtype: str, count: int) -> np.ndarray:
        dt = np.dtype(self.e + dtype)
        return np.frombuffer(self.take(dt.itemsize * count), dtype=dt).astype(dt.newbyteorder("="))


def _lm_max(l_max: int, m_max: int, minc: int) -> int:
    return sum(l_max - m + 1 for m in range(0, m_max + 1, minc))


def read_checkpoint(path) -> CheckpointData:
    """Read a MagIC checkpoint (``checkpoint_io.py:100-309``): the whole file,
    checked to end exactly after the last field."""
    with open(path, "rb") as f:
        buf = f.read()
    if len(buf) < 4:
        raise ValueError("checkpoint too short")
    v_le, v_be = struct.unpack("<i", buf[:4])[0], struct.unpack(">i", buf[:4])[0]
    if 2 <= v_le <= 5:
        e, ver = "<", v_le
    elif 2 <= v_be <= 5:
        e, ver = ">", v_be
    else:
        raise ValueError(f"cannot detect endianness: version word reads {v_le} (LE) / {v_be} (BE)")
    rd = _Reader(buf, e)
    rd.take(4)
    ck = CheckpointData(version=ver, endian=e)
    (ck.time,) = rd.unpack("d")
    ck.family = rd.take(10).decode("ascii").strip()
    ck.nexp, ck.nimp, ck.nold = rd.unpack("3i")
    ck.dt = rd.array("f8", 1 if ck.family == "DIRK" else ck.nexp)
    (ck.n_time_step,) = rd.unpack("i")
    if ver >= 3:
        (ck.ra, ck.pr, ck.raxi, ck.sc, ck.prmag, ck.ek, ck.stef, ck.radratio,
         ck.sigma_ratio) = rd.unpack("9d")
    else:
        (ck.ra, ck.pr, ck.raxi, ck.sc, ck.prmag, ck.ek, ck.radratio,
         ck.sigma_ratio) = rd.unpack("8d")
    (ck.n_r_max, ck.n_theta_max, ck.n_phi_tot, ck.minc, ck.nalias,
     ck.n_r_ic_max) = rd.unpack("6i")
    if ver >= 4:
        ck.l_max, ck.m_min, ck.m_max = rd.unpack("3i")
    else:
        ck.l_max = ck.nalias * ck.n_phi_tot // 60
        ck.m_min, ck.m_max = 0, ck.l_max
    rd.take(72)                                   # radial scheme name
    ck.n_cheb_max, _ = rd.unpack("2i")
    rd.unpack("2d")                               # mapping parameters alph1, alph2
    ck.r = rd.array("f8", ck.n_r_max)
    nh = lambda k: max(0, k - 1)
    if ck.family == "MULTISTEP":
        def scalars():
            return {part: [rd.unpack("d")[0] for _ in range(nh(k))]
                    for part, k in (("expl", ck.nexp), ("impl", ck.nimp), ("old", ck.nold))}
        ck.domega_ic_dt = scalars()
        ck.domega_ma_dt = scalars()
        if ver < 5:   # Lorentz-torque histories, not kept
            rd.take(16 * (nh(ck.nexp) + nh(ck.nimp) + nh(ck.nold)))
    om = rd.unpack("12d")
    ck.omega_ic1, ck.omega_ma1 = om[0], om[6]
    if ver == 2:
        fl = rd.unpack("5i")
        ck.l_heat, ck.l_chemical_conv, ck.l_mag, ck.l_press_store, ck.l_cond_ic = map(bool, fl)
    else:
        fl = rd.unpack("6i")
        (ck.l_heat, ck.l_chemical_conv, ck.l_phase_field, ck.l_mag, ck.l_press_store,
         ck.l_cond_ic) = map(bool, fl)
    L = _lm_max(ck.l_max, ck.m_max, ck.minc)

    def arr(n_r):
        return rd.array("c16", L * n_r).reshape((L, n_r), order="F")

    def one(name, n_r):
        a = arr(n_r)
        if ck.family == "MULTISTEP":
            ck.derivatives[name] = {part: [arr(n_r) for _ in range(nh(k))]
                                    for part, k in (("expl", ck.nexp), ("impl", ck.nimp),
                                                    ("old", ck.nold))}
        return a

    ck.w = one("w", ck.n_r_max)
    ck.z = one("z", ck.n_r_max)
    if ck.l_press_store:
        ck.p = one("p", ck.n_r_max)
    if ck.l_heat:
        ck.s = one("s", ck.n_r_max)
    if ck.l_chemical_conv:
        ck.xi = one("xi", ck.n_r_max)
    if ck.l_phase_field:
        ck.phi = one("phi", ck.n_r_max)
    if ck.l_mag:
        ck.b = one("b", ck.n_r_max)
        ck.aj = one("aj", ck.n_r_max)
        if ck.l_cond_ic:
            ck.b_ic = one("b_ic", ck.n_r_ic_max)
            ck.aj_ic = one("aj_ic", ck.n_r_ic_max)
    if rd.pos != len(buf):
        raise ValueError(f"checkpoint has {len(buf) - rd.pos} bytes after the last field")
    return ck


# CNAB2 slot counts and the order the fields are written in
_NEXP, _NIMP, _NOLD = 2, 1, 1
_SYS_OF = {"w": "w", "z": "z", "p": "p", "s": "s", "b": "b", "aj": "j"}
_IC_SYS_OF = {"b_ic": "b", "aj_ic": "j"}


def _history_slot(run, sys_key):
    """The explicit term the next step reads as history (the reference's
    expl[:, :, 1] after a whole step)."""
    return run.expl_cur[sys_key] if getattr(run, "expl_pending", False) else \
        run.expl_prev[sys_key]


def _ic_history_slot(run, sys_key):
    """The inner core's explicit history (expl[:, :, 1] after a whole step)."""
    return run.expl_ic[1 - run._par][sys_key]


def write_checkpoint_fortran(path, run, sim_time: float, n_time_step: int,
                             endian: str = ">") -> None:
    """Write `run`'s state as a MagIC v5 checkpoint (``checkpoint_io.py:336-490``;
    time scale 1 as in the Boussinesq benchmarks, ``pre_calculations.py:13-14``):
    MULTISTEP (CNAB2) with the explicit histories of the fields and of the inner
    core's rotation, or DIRK (BPR353) with the fields only; a conducting inner
    core adds its fields b_ic, aj_ic (and their histories)."""
    p = run.p
    dirk = getattr(run, "dirk", False)
    cond_ic = getattr(run, "cond_ic", False)
    if getattr(run, "expl_pending", False):
        raise RuntimeError("write_checkpoint_fortran: a radial loop awaits its lm loop; "
                           "write after a whole step")
    e = endian
    out = bytearray()
    pk = lambda fmt, *v: out.extend(struct.pack(e + fmt, *v))
    pk("i", 5)
    pk("d", float(sim_time))
    if dirk:
        out.extend("DIRK".ljust(10).encode("ascii"))
        pk("3i", 4, 4, 1)
        pk("d", float(run.dt_hist[0]))
    else:
        out.extend("MULTISTEP".ljust(10).encode("ascii"))
        pk("3i", _NEXP, _NIMP, _NOLD)
        pk("2d", float(run.dt_hist[0]), float(run.dt_hist[1]))
    pk("i", int(n_time_step))
    pk("9d", p.ra, p.pr, p.raxi, p.sc, p.prmag, p.ek, p.stef, p.radratio, p.sigma_ratio)
    pk("6i", p.n_r_max, p.n_theta, p.n_phi_tot, p.minc, 20, (p.n_r_max + 1) // 2)
    pk("3i", p.l_max, 0, p.l_max)
    out.extend("cheb".ljust(72).encode("ascii"))
    pk("2i", p.n_cheb_max, 0)
    pk("2d", p.alph1, p.alph2)
    out.extend(np.asarray(run.res.rad["r"], dtype=e + "f8").tobytes())
    if not dirk:
        # rotation-rate histories: expl slot 1 of omega_ic (the Lorentz-torque
        # term of the last step), of omega_ma (not evolved here)
        om_hist = float(run.rot[7 - run._par].item()) if cond_ic else 0.0
        pk("d", om_hist)
        pk("d", 0.0)
    pk("12d", p.omega_ic1, *([0.0] * 11))
    for flag in (p.l_heat, False, False, p.l_mag, True, p.l_cond_ic):
        pk("i", 1 if flag else 0)

    def field_bytes(t):
        a = t.detach().to("cpu", torch.complex128).numpy()
        return np.asarray(a, dtype=e + "c16").ravel(order="F").tobytes()

    names = ["w", "z", "p"] + (["s"] if p.l_heat else []) + (["b", "aj"] if p.l_mag else [])
    for name in names:
        out.extend(field_bytes(run.f[name]))
        if not dirk:
            out.extend(field_bytes(_history_slot(run, _SYS_OF[name])))
    if cond_ic:
        for name, key in _IC_SYS_OF.items():
            out.extend(field_bytes(run.fic[name]))
            if not dirk:
                out.extend(field_bytes(_ic_history_slot(run, key)))
    with open(path, "wb") as f:
        f.write(bytes(out))


def _translate_s00(run):
    """Shift the l=0 entropy profile to s = 0 at the outer wall when it holds
    the legacy conduction values (reference ``main.py:82-99``)."""
    rad = run.res.rad
    r_cmb, r_icb = rad["r_cmb"], rad["r_icb"]
    sq4pi = math.sqrt(4.0 * math.pi)
    top = -r_icb ** 2 / (r_icb ** 2 + r_cmb ** 2) * sq4pi
    bot = r_cmb ** 2 / (r_icb ** 2 + r_cmb ** 2) * sq4pi
    eps = np.finfo(np.float64).eps
    s00 = run.f["s"][0].cpu().numpy()
    if abs(s00[0].real - top) <= 10000 * eps and abs(s00[-1].real - bot) <= 10000 * eps:
        run.f["s"][0] -= top


def load_fortran_checkpoint(run, path) -> CheckpointData:
    """Fields of a MagIC checkpoint into `run` (reference
    ``main.load_fortran_checkpoint``, ``main.py:121-166``); returns the data."""
    ck = read_checkpoint(path)
    if ck.minc != run.p.minc:
        raise ValueError(f"checkpoint has minc={ck.minc}, the run has minc={run.p.minc}")
    shape = tuple(run.f["w"].shape)
    for name in ("w", "z", "p", "s", "b", "aj"):
        a = getattr(ck, name)
        if a is None:
            continue
        if a.shape != shape:
            raise ValueError(f"checkpoint field {name} is {a.shape}, the run is {shape}")
        if name in ("b", "aj") and not run.p.l_mag:
            continue
        run.f[name].copy_(torch.from_numpy(np.ascontiguousarray(a)))
        if name == "s" and run.p.l_heat:
            _translate_s00(run)
    if ck.xi is not None and np.any(ck.xi):
        raise ValueError("checkpoint holds a composition field; this build does not evolve it")
    # inner-core fields when the run has a conducting core (kept otherwise, as the
    # reference does); omega_ic is set from z(l=1, m=0) by setup_initial_state
    if getattr(run, "cond_ic", False):
        for name in ("b_ic", "aj_ic"):
            a = getattr(ck, name)
            if a is None:
                continue
            if a.shape != tuple(run.fic[name].shape):
                raise ValueError(f"checkpoint field {name} is {a.shape}, the run is "
                                 f"{tuple(run.fic[name].shape)}")
            run.fic[name].copy_(torch.from_numpy(np.ascontiguousarray(a)))
    return ck


def restart_from_fortran(run, path) -> tuple[float, float]:
    """The fresh-start sequence of ``main.run`` for a Fortran checkpoint
    (``main.py:318-374``): fields from the file, ``setup_initial_state``, dt
    from the file through ``initialize_dt``, then the file's explicit history
    restored into the slot the next step reads.  Returns (sim_time, dt); the
    run continues with ``one_step(1, dt)`` (step numbers restart at 1, so the
    first step rebuilds the operators, as in the reference)."""
    ck = load_fortran_checkpoint(run, path)
    run.setup_initial_state()
    dt = float(ck.dt[0]) if len(ck.dt) else run.p.dtmax
    run.initialize_dt(dt)
    if getattr(run, "dirk", False):
        # a DIRK file holds no histories, and a BPR353 step reads none
        run.time = ck.time
        return ck.time, run.dt_hist[0]
    cond_ic = getattr(run, "cond_ic", False)
    for name, d in ck.derivatives.items():
        key = _SYS_OF.get(name)
        if key is not None and (key not in ("b", "j") or run.p.l_mag):
            for i, a in enumerate(d["expl"][:1]):
                run.expl_prev[key].copy_(torch.from_numpy(np.ascontiguousarray(a)))
        key = _IC_SYS_OF.get(name)
        if key is not None and cond_ic:
            for i, a in enumerate(d["expl"][:1]):
                _ic_history_slot(run, key).copy_(torch.from_numpy(np.ascontiguousarray(a)))
    if getattr(run, "rot_ic", False) and ck.domega_ic_dt and ck.domega_ic_dt["expl"]:
        run.rot[7 - run._par] = float(ck.domega_ic_dt["expl"][0])
    run.time = ck.time
    return ck.time, run.dt_hist[0]
