"""ctypes binding of ``libmagic_hip.so`` (the C ABI declared in include/magic_hip.h).

The structures below mirror the header field by field; :func:`load` checks
their sizes against ``mh_struct_sizes`` so a layout drift fails loudly.
There is no fallback: when the library is missing every product entry point
raises :class:`MagicHipError`.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmagic_hip.so")

P = C.c_void_p
I = C.c_int
L = C.c_long
D = C.c_double


class MagicHipError(RuntimeError):
    """Raised when the HIP library is missing or a kernel entry point fails."""


class SynIn(C.Structure):
    _fields_ = [("x1", P), ("x2", P), ("mode", I), ("pad", I)]


class SynArgs(C.Structure):
    _fields_ = [("in_", (SynIn * 3) * 4), ("ngroups", I), ("B", I), ("ldx", I), ("H", I),
                ("Hp", I), ("mp_info", P), ("lm_start", P), ("synP", P), ("synD", P),
                ("or2", P), ("F", P), ("c_stride", L), ("m_stride", L), ("Bout", I), ("prec", I), ("fprec", I), ("waves", I), ("nct", I), ("minc", I),
                ("mp16", P)]


class AnaArgs(C.Structure):
    _fields_ = [("F", P), ("c_stride", L), ("m_stride", L), ("Bin", I), ("B", I), ("ncomp", I),
                ("cf1", I * 3), ("cf2", I * 3), ("pad0", I), ("outS", P * 3), ("outT", P * 3),
                ("ldo", I), ("H", I), ("Hp", I), ("n_m", I), ("mp_info", P), ("lm_start", P),
                ("anaW", P), ("anaWd", P), ("tiles", P), ("ntiles", I), ("prec", I), ("fprec", I), ("waves", I),
                ("minc", I), ("mp16", P)]


class FftArgs(C.Structure):
    _fields_ = [("Mh", I), ("nstage", I), ("radix", I * 16), ("n_m", I), ("T", I), ("B", I),
                ("ncomp", I), ("RB", I), ("pad", I), ("tw", P), ("tw2", P), ("F", P),
                ("Fc_stride", L), ("Fm_stride", L), ("FB", I), ("pad0", I), ("G", P),
                ("Gc_stride", L), ("comp_scale", P), ("ost2", P)]


class NlArgs(C.Structure):
    _fields_ = [("Gs", P), ("Gv", P), ("Gv_c_stride", L), ("Go", P), ("Go_c_stride", L),
                ("N", I), ("T", I), ("P", I), ("pad", I), ("or2", P), ("or4", P), ("ost2", P),
                ("valri2", P), ("valhi2", P), ("LFfac", D), ("cf2", D), ("af2", D),
                ("l_mag", I), ("alf_damp", I), ("cfl", P)]


class GridArgs(C.Structure):
    _fields_ = [("Fs", P), ("Fv", P), ("Fo", P), ("N", I), ("T", I), ("M", I), ("Mh", I),
                ("generic", I), ("fprec", I), ("tw", P), ("tw2", P), ("or2", P),
                ("or4", P), ("ost2", P), ("valri2", P), ("valhi2", P), ("LFfac", D), ("cf2", D),
                ("af2", D), ("l_mag", I), ("alf_damp", I), ("cfl", P), ("rb", I), ("pad", I)]


class RowJob(C.Structure):
    _fields_ = [("X", P), ("rows", P), ("MT", P), ("Y", P * 6), ("ldx", I), ("ldy", I),
                ("K", I), ("R", I), ("Rpad", I), ("rseg", I), ("nrows", I), ("pad", I)]


class LmCoef(C.Structure):
    _fields_ = [("l", P), ("m", P), ("dLh", P), ("dTh2A", P), ("dTh2S", P), ("dTh3A", P),
                ("dTh3S", P), ("dTh4A", P), ("dTh4S", P), ("l_max", I), ("L", I)]


class RadCoef(C.Structure):
    _fields_ = [("or1", P), ("or2", P), ("or3", P), ("or4", P), ("r2", P), ("rgrav", P),
                ("N", I), ("pad", I)]


class TdArgs(C.Structure):
    _fields_ = [("lc", LmCoef), ("rc", RadCoef), ("Advr", P), ("Advt", P), ("Advp", P),
                ("VSt", P), ("VxBr", P), ("VxBp", P), ("dVSr_dr", P), ("dVxBt_dr", P),
                ("w", P), ("dw", P), ("z", P), ("ew", P), ("ez", P), ("ep", P), ("es", P),
                ("eb", P), ("ej", P), ("CorFac", D), ("l_mag", I), ("pad", I)]


class ImexSys(C.Structure):
    _fields_ = [("old", P), ("impl", P), ("expl0", P), ("expl1", P),
                ("t4", P), ("t5", P), ("t6", P), ("t7", P)]


class RhsArgs(C.Structure):
    _fields_ = [("sys", ImexSys * 6), ("rs", P), ("rz", P), ("rb", P), ("rj", P), ("rwp", P),
                ("m", P), ("L", I), ("N", I), ("wimp", D), ("wl1", D), ("we0", D), ("we1", D),
                ("bots_l0", D), ("l_mag", I), ("pad", I), ("wdev", P)]


class IcRhsArgs(C.Structure):
    _fields_ = [("sys", ImexSys * 2), ("rb", P), ("rj", P), ("out_b", P), ("out_j", P),
                ("m", P), ("L", I), ("N", I), ("n_ic", I), ("pad", I), ("wimp", D), ("wl1", D),
                ("we0", D), ("we1", D), ("wdev", P)]


class IcImplArgs(C.Structure):
    _fields_ = [("b", P), ("db", P), ("ddb", P), ("aj", P), ("dj", P), ("ddj", P),
                ("old_b", P), ("old_j", P), ("impl_b", P), ("impl_j", P), ("dLh", P), ("l", P),
                ("O_r_ic", P), ("L", I), ("n_ic", I), ("fac", D), ("or2_icb", D)]


class IcRotArgs(C.Structure):
    _fields_ = [("rot", P), ("wdev", P), ("term", I * 8), ("expl_slot", I), ("impl_slot", I),
                ("store_old", I), ("lm10", I), ("L", I), ("N", I), ("n_ic", I), ("T", I),
                ("P", I), ("c_z10_omega", D), ("c_dt_z10", D), ("c_lorentz", D), ("fac_lt", D),
                ("or1_icb", D), ("or2_icb", D), ("or4_icb", D), ("r_icb", D), ("br", P),
                ("bp", P), ("vt", P), ("vp", P), ("gauss", P), ("sin2", P), ("b_ic", P),
                ("aj_ic", P), ("e_b", P), ("e_j", P), ("m", P), ("dLh", P), ("rhs_z", P),
                ("z", P), ("dz", P)]


class BuoyArgs(C.Structure):
    _fields_ = [("rwp", P), ("s", P), ("ew0", P), ("rgrav", P), ("m", P), ("L", I), ("N", I),
                ("fac", D), ("BuoFac", D), ("lm_l0", I), ("pad", I), ("wdev", P)]


class ImplArgs(C.Structure):
    _fields_ = [("lc", LmCoef), ("rc", RadCoef)] + [(n, P) for n in (
        "s", "ds", "d2s", "z", "d2z", "w", "dw", "ddw", "dddw", "p", "dp", "b", "ddb", "aj", "ddj",
        "old_s", "impl_s", "old_z", "impl_z", "old_w", "impl_w", "old_p", "impl_p",
        "old_b", "impl_b", "old_j", "impl_j")] + [("opr", D), ("opm", D), ("BuoFac", D),
                                                  ("l_mag", I), ("lm_l0", I)]


class EnergyArgs(C.Structure):
    _fields_ = [("w", P), ("dw", P), ("z", P), ("b", P), ("db", P), ("aj", P), ("lm_start", P),
                ("or2", P), ("qw", P), ("out", P), ("l_max", I), ("N", I), ("by_m", I),
                ("l_mag", I), ("LFfac", D), ("minc", I), ("pad", I), ("orho1", P)]


class AnelProf(C.Structure):
    _fields_ = [(n, P) for n in ("beta", "dbeta", "rho0", "orho1", "otemp1", "dLtemp0", "visc",
                                 "dLvisc", "kappa", "dLkappa", "rgrav_rho")] + [
        ("ViscHeatFac", D), ("OhmLossFac", D)]


class NlAnelArgs(C.Structure):
    _fields_ = [("Gs", P), ("Gv", P), ("Gv_c_stride", L), ("Gb", P), ("Gb_c_stride", L),
                ("Gd", P), ("Gd_c_stride", L), ("Go", P), ("Go_c_stride", L), ("N", I), ("T", I),
                ("P", I), ("l_mag", I), ("or1", P), ("or2", P), ("or4", P), ("ost2", P),
                ("cosn2", P), ("valri2", P), ("valhi2", P), ("prof", AnelProf), ("LFfac", D),
                ("cf2", D), ("af2", D), ("alf_damp", I), ("pad", I), ("cfl", P)]


class DphiArgs(C.Structure):
    _fields_ = [("F", P), ("Fc_stride", L), ("Fm_stride", L), ("FB", I), ("c0", I), ("out", P),
                ("Oc_stride", L), ("Om_stride", L), ("OB", I), ("B", I), ("n_m", I), ("T", I),
                ("minc", I), ("pad", I)]


class EntropyAnelArgs(C.Structure):
    _fields_ = [("es", P), ("heat", P), ("orho1", P), ("L", I), ("N", I)]


class AmArgs(C.Structure):
    _fields_ = [("z", P), ("dz", P), ("d2z", P), ("r2q", P), ("pz", P), ("pdz", P), ("pd2z", P),
                ("l1m0", I), ("l1m1", I), ("N", I), ("do_z", I), ("do_e", I), ("pad", I),
                ("cz", D), ("ce", D)]


class P0IntArgs(C.Structure):
    _fields_ = [("rwp", P), ("s", P), ("u", P), ("lm_l0", I), ("N", I)]


class ImplAnelArgs(C.Structure):
    _fields_ = [("base", ImplArgs), ("prof", AnelProf), ("dz", P), ("zero_edges", I), ("pad", I)]


class BuildArgs(C.Structure):
    _fields_ = [("R", P), ("D1", P), ("D2", P), ("D3", P), ("or1", P), ("or2", P), ("out", P),
                ("N", I), ("l_max", I), ("n_cheb", I), ("ktopv", I), ("kbotv", I), ("kind", I),
                ("rnorm", D), ("wimp", D), ("opr", D), ("opm", D), ("wimp_dev", P), ("when", P),
                ("out_wp", P), ("out_p0", P), ("nsys", I), ("pad", I)]


class GemmArgs(C.Structure):
    _fields_ = [("A", P), ("B", P), ("C", P)] + [(n, L) for n in (
        "sAr", "sAc", "sAb", "sBr", "sBc", "sBb", "sCr", "sCc", "sCb")] + [
        ("M", I), ("Nn", I), ("K", I), ("batch", I), ("alpha", D), ("when", P)]


class StepState(C.Structure):
    _fields_ = [("dt", D), ("dtMax", D), ("dt1", D), ("dt2", D), ("wimp_lin0", D), ("time", D),
                ("n_step", I), ("rebuild", I), ("n_rebuild", I), ("new_dt", I),
                ("sing_step", I), ("sing_sys", I), ("sing_l", I), ("sing_info", I)]


class StepCtlArgs(C.Structure):
    _fields_ = [("cfl", P), ("delxr2", P), ("delxh2", P), ("state", P), ("w", P),
                ("dt_schedule", P), ("dt_log", P), ("info", P), ("N", I), ("n_schedule", I),
                ("log_cap", I), ("info_rows", I), ("info_cols", I), ("pad", I), ("alpha", D),
                ("BuoFac", D), ("w_stage", P), ("scheme", I), ("pad2", I)]


class DcArgs(C.Structure):
    _fields_ = [("lc", LmCoef), ("rc", RadCoef), ("beta", P), ("dbeta", P), ("ddbeta", P),
                ("orho1", P), ("rho0", P), ("visc", P), ("dLvisc", P), ("ddLvisc", P),
                ("grav", P), ("Advr", P), ("dAdvt_dr", P), ("w", P), ("dw", P), ("ddw", P),
                ("dddw", P), ("ddddw", P), ("z", P), ("dz", P), ("s", P), ("ew", P), ("ep", P),
                ("rwp", P), ("ew0", P), ("wdev", P), ("old_w", P), ("impl_w", P),
                ("old_p", P), ("impl_p", P), ("CorFac", D), ("BuoFac", D), ("fac", D),
                ("lm_l0", I), ("pad", I)]


_STRUCT_ORDER = [SynArgs, AnaArgs, FftArgs, NlArgs, RowJob, TdArgs, RhsArgs, BuoyArgs,
                 ImplArgs, EnergyArgs, BuildArgs, GemmArgs, GridArgs, StepState, StepCtlArgs,
                 IcRhsArgs, IcImplArgs, IcRotArgs, NlAnelArgs, DphiArgs, EntropyAnelArgs,
                 ImplAnelArgs, AmArgs, P0IntArgs, DcArgs]

_SIGS = {
    "mh_plm_tables": [I, I, I, I, I, P, P, P, P, P, P, P, P, P],
    "mh_leg_synth": [I, P, I, P],
    "mh_leg_analysis": [I, P, P],
    "mh_fft": [I, P, P],
    "mh_get_nl": [P, P],
    "mh_nl_anel": [P, P],
    "mh_fourier_dphi": [P, P],
    "mh_entropy_anel": [P, P],
    "mh_impl_anel": [P, P],
    "mh_am_correct": [P, P],
    "mh_p0_integral": [P, P],
    "mh_grid_step": [P, P],
    "mh_grid_plan": [I, P],
    "mh_ic_rhs": [P, P],
    "mh_ic_impl": [P, P],
    "mh_ic_torque": [P, P],
    "mh_ic_advect": [P, P],
    "mh_ic_omega_rhs": [P, P],
    "mh_ic_omega_update": [P, P],
    "mh_tables_to_f32": [P, P, L, P],
    "mh_tables_to_bf16": [P, P, P, P, P, P, I, I, I, I, P, P, P, P, P],
    "mh_rowop": [P, P, I, I, P],
    "mh_td": [P, P],
    "mh_rhs": [P, P],
    "mh_buoy": [P, P],
    "mh_impl": [P, P],
    "mh_energy": [P, P],
    "mh_build_matrices": [P, P],
    "mh_gj_inverse": [P, P, P, P, I, I, I, P, P, P],
    "mh_gj_inverse_noalloc": [P, P, I, I, I, P, P, P],
    "mh_gemm_f64": [P, P],
    "mh_step_control": [P, P],
    "mh_fill_f64": [P, L, I, L, D, P, P],
    "mh_band_factor": [P, P, P, I, I, I, I, P],
    "mh_band_solve": [P, P, P, P, P, P, P, P, I, I, I, I, P],
    "mh_band_factor_gated": [P, P, P, I, I, I, I, P, P],
    "mh_band_solve_gated": [P, P, P, P, P, P, P, P, I, I, I, I, P, P],
    "mh_td_dc": [P, P],
    "mh_buoy_dc": [P, P],
    "mh_impl_dc": [P, P],
    "mh_affine_f64": [P, P, P, L, D, P, P, P],
    "mh_struct_sizes": [P, I],
}

EXPORTED = tuple(_SIGS)

GJ_SCRATCH_DOUBLES = 4 << 20   # MH_GJ_SCRATCH_DOUBLES
_ERRORS = {-1: "invalid argument", -2: "launch failure",
           -3: f"sizes exceed the fixed work space ({GJ_SCRATCH_DOUBLES} doubles: "
               "nbatch * (2n^2 + 2n) must fit)"}

_lib = None


def load(path: str | None = None):
    """Load the library (once), bind signatures and verify struct sizes."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise MagicHipError(f"HIP library not built: {path} (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = I
    buf = (I * 32)()
    n = lib.mh_struct_sizes(C.cast(buf, P), 32)
    if n != len(_STRUCT_ORDER):
        raise MagicHipError(f"struct table mismatch: library reports {n} structs")
    for k, cls in enumerate(_STRUCT_ORDER):
        if buf[k] != C.sizeof(cls):
            raise MagicHipError(f"{cls.__name__}: binding size {C.sizeof(cls)} != library {buf[k]}")
    _lib = lib
    return lib


_tls = threading.local()  # per thread: the launch plan being recorded (or None)


def recording():
    """The launch plan this thread is recording, or None."""
    return getattr(_tls, "rec", None)


class record:
    """Collect the entry-point calls made inside the block (in this thread)
    instead of running them: ``with record() as plan: ...``; each item is
    (section, name, args without the trailing stream).  Every entry point
    takes the stream last."""

    def __init__(self):
        self.calls = []
        self.section = None

    def __enter__(self):
        if recording() is not None:
            raise RuntimeError("nested launch-plan recording")
        _tls.rec = self
        return self

    def __exit__(self, *exc):
        _tls.rec = None
        return False


def call(name: str, *args):
    """Call an entry point and raise on a non-zero return code."""
    rec = recording()
    if rec is not None:
        rec.calls.append((rec.section, name, args[:-1]))
        return 0
    rc = getattr(load(), name)(*args)
    if rc != 0:
        raise MagicHipError(f"{name}: {_ERRORS.get(rc, f'code {rc}')}")
    return rc


def ref(struct) -> C.c_void_p:
    """Host pointer to a ctypes structure, for the ``const Mh*Args*`` parameters."""
    return C.cast(C.pointer(struct), P)


def run_plan(calls, stream, on_section=None):
    """Run recorded calls on `stream` (a c_void_p); on_section(name, begin) brackets
    each run of consecutive calls of one section (for timers)."""
    lib = load()
    cur = None
    for sec, name, args in calls:
        if on_section is not None and sec != cur:
            if cur is not None:
                on_section(cur, False)
            on_section(sec, True)
            cur = sec
        rc = getattr(lib, name)(*args, stream)
        if rc != 0:
            raise MagicHipError(f"{name}: {_ERRORS.get(rc, f'code {rc}')}")
    if on_section is not None and cur is not None:
        on_section(cur, False)
