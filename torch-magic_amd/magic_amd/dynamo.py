"""One dynamo run resident on the GPU: state, implicit operators and the CNAB2 step.

:class:`DynamoRun` is the context object behind the reference's module-level
state (``fields.py``, ``dt_fields.py``, the per-l matrices of ``update_*.py``)
and step driver (``step_time.py``).  All per-step arithmetic runs in
libmagic_hip kernels on one HIP stream; the host only launches kernels, reads
back the 2*n_r_max Courant maxima once per step (the reference's
``dtrkc.min().item()`` sync, ``step_time.py:1212``) and, when dt changes,
re-runs the on-device operator rebuild.

Per step (``_one_step_cnab2``, reference ``step_time.py:1602-1637``):

radial loop (``step_time.py:931-1212``)
  1. Legendre synthesis of s and of the QST groups (B, curl B, u, curl u)
     into Fourier buffers (MFMA FP64), the l(l+1) and or2 prefactors applied
     while loading;
  2. inverse azimuthal FFT to the grid;
  3. nonlinear products + per-level Courant maxima;
  4. forward FFT of the 9 product components of the bulk levels;
  5. Legendre analysis (scalar and spheroidal/toroidal);
  6. radial derivatives of the two flux terms (row-operator GEMM);
  7. explicit terms with Coriolis coupling.
dt decision (``courant.py:130-155``), IMEX weights (``time_scheme.py:56-73``),
operator rebuild if dt changed (``step_time.py:1530-1548``);
lm loop (``step_time.py:222-394``, ``update_b.py:557-631``)
  8. IMEX right-hand sides with boundary rows;
  9. per-degree solves of S, Z, B, J with the DCT-I and the first two radial
     derivatives folded into one operator per degree (row-operator GEMM);
 10. buoyancy coupling and the l=0 pressure right-hand side;
 11. per-degree (w, p) solve with w, dw, ddw, dddw, p, dp folded in;
 12. stored old states and implicit terms; explicit-history rotation is a
     pointer swap.
"""

from __future__ import annotations

import ctypes as C
import gc
import math

import numpy as np
import torch

from . import _lib, grid
from .device import C128, F64, I32, Resolution, ptr, require_gpu, stream_handle
from .params import LEGENDRE_PREC, Params

_FIELDS = ("s", "ds", "z", "dz", "w", "dw", "ddw", "p", "dp", "b", "db", "ddb", "aj", "dj", "ddj")
_SYS = ("s", "z", "w", "p", "b", "j")


def _rowjob_dtype():
    """numpy record layout of MhRowJob (from the ctypes mirror's field offsets)."""
    names, formats, offsets = [], [], []
    for name, ct in _lib.RowJob._fields_:
        names.append(name)
        offsets.append(getattr(_lib.RowJob, name).offset)
        formats.append((np.uint64, 6) if name == "Y" else
                       (np.uint64 if ct is _lib.P else np.int32))
    return np.dtype(dict(names=names, formats=formats, offsets=offsets,
                         itemsize=C.sizeof(_lib.RowJob)))


_ROWJOB_DTYPE = _rowjob_dtype()


def pack_row_jobs(specs, tile_cols=64, n_xcd=8):
    """Pack row-operator job columns into MhRowJob records and build the tile
    list (job, first row, first column): rows in chunks of 32, operator columns
    in chunks of `tile_cols` (64, or 128 for the wide kernel).  Each spec is a
    dict of MhRowJob fields whose rows / nrows / MT entries may be arrays (one
    job per element).

    The hardware hands workgroup i to XCD i mod 8, and each XCD has its own L2.
    So that the operator of a job is read from HBM by one XCD only (and its
    right-hand-side rows by one XCD only), the tiles are ordered XCD by XCD
    (`xcd_tile_order`); `n_xcd=1` keeps plain job order."""
    n = [int(np.size(sp["nrows"])) for sp in specs]
    tot = sum(n)
    arr = np.zeros(tot, dtype=_ROWJOB_DTYPE)
    k = 0
    for sp, m in zip(specs, n):
        sl = arr[k:k + m]
        for name in ("X", "rows", "MT", "ldx", "ldy", "K", "R", "Rpad", "rseg", "nrows"):
            sl[name] = sp[name]
        sl["Y"] = np.asarray(list(sp["Y"]) + [0] * (6 - len(sp["Y"])), np.uint64)[None, :]
        k += m
    nr = -(-arr["nrows"].astype(np.int64) // 32)
    nc = -(-arr["R"].astype(np.int64) // tile_cols)
    per = nr * nc
    ji = np.repeat(np.arange(tot), per)
    q = np.arange(per.sum()) - np.repeat(np.cumsum(per) - per, per)
    ncj = np.repeat(nc, per)
    tiles = np.stack([ji, 32 * (q // ncj), tile_cols * (q % ncj)], axis=1).astype(np.int32)
    if n_xcd > 1 and len(tiles):
        tiles = tiles[xcd_tile_order(per, nc, n_xcd)]
    return arr, tiles


def xcd_tile_order(per, nc, n_xcd=8):
    """Permutation of a job-ordered tile list (job j owns `per[j]` consecutive
    tiles, `nc[j]` per 32-row chunk) that gives each unit of work to one XCD.

    Units are whole jobs, except that a job holding more than 1/64 of all tiles
    is cut at row-chunk boundaries into pieces of about that size.  Units are
    dealt largest first to the least-loaded of `n_xcd` queues (ties: the lower
    queue), each unit's tiles staying consecutive in its queue; the queues are
    then interleaved, position i of the result taking the next tile of queue
    i mod n_xcd while every queue has tiles left.  The few tiles left when the
    queues' lengths differ follow at the end, in queue order."""
    per = np.asarray(per, np.int64)
    nc = np.asarray(nc, np.int64)
    total = int(per.sum())
    start = np.cumsum(per) - per
    cap = max(1, total // 64)
    units = []  # (size, first tile, job)
    for j in np.nonzero(per)[0]:
        step = max(1, cap // int(nc[j])) * int(nc[j])
        for s0 in range(0, int(per[j]), step):
            units.append((min(step, int(per[j]) - s0), int(start[j]) + s0))
    order = sorted(range(len(units)), key=lambda u: (-units[u][0], u))
    load = [0] * n_xcd
    queues = [[] for _ in range(n_xcd)]
    for u in order:
        x = min(range(n_xcd), key=lambda i: (load[i], i))
        size, s0 = units[u]
        queues[x].append(np.arange(s0, s0 + size))
        load[x] += size
    qs = [np.concatenate(qx) if qx else np.zeros(0, np.int64) for qx in queues]
    m = min(len(qx) for qx in qs)
    head = np.stack([qx[:m] for qx in qs], axis=1).reshape(-1)
    tail = [qx[m:] for qx in qs]
    return np.concatenate([head] + tail).astype(np.int64)


def seeded_noise(p: Params, res: Resolution, seed: int, amp: float) -> np.ndarray:
    """amp * complex N(0,1) per lm (l >= 1, m = 0 real) times (1 - x^2)^3 in r.

    Drawn on the CPU in FP64 from ``numpy.random.default_rng(seed)`` (real
    parts for all lm, then imaginary parts) in m-major lm order; the radial
    shape vanishes at both boundaries so the fixed-entropy rows hold.
    """
    coef, shape = noise_factors(p, res, seed, amp)
    return coef[:, None] * shape[None, :]


def noise_factors(p: Params, res: Resolution, seed: int, amp: float):
    """The seeded noise as its two factors: (coefficient per lm, radial shape)."""
    rng = np.random.default_rng(seed)
    L = p.lm_max
    coef = amp * (rng.standard_normal(L) + 1j * rng.standard_normal(L))
    m, l = res.maps["m"], res.maps["l"]
    coef[m == 0] = coef[m == 0].real
    coef[l == 0] = 0.0
    r = res.rad["r"]
    x = 2.0 * r - res.rad["r_cmb"] - res.rad["r_icb"]
    return coef, (1.0 - x * x) ** 3


def conduction_state(p: Params, res: Resolution):
    """Boussinesq conduction profile (s0, p0) on the radial grid.

    Solves the reference's coupled Chebyshev collocation system
    (``init_fields.py:108-180``): opr (s'' + 2/r s') = 0, p' = BuoFac rgrav s,
    s(r_cmb) = 0, s(r_icb) = sqrt(4 pi), p(r_cmb) = 0, with the boundary
    factor 1/2 on the first/last Chebyshev columns; host LAPACK, run once.
    """
    rad = res.rad
    N = p.n_r_max
    rn = rad["rnorm"]
    T, T1, T2 = rad["T"], rad["T1"], rad["T2"]
    A = np.zeros((2 * N, 2 * N))
    A[:N, :N] = rn * p.opr * (T2 + (2.0 * rad["or1"])[:, None] * T1)
    A[N:, :N] = -rn * (p.BuoFac * rad["rgrav"])[:, None] * T
    A[N:, N:] = rn * T1
    A[0, :] = 0.0; A[0, :N] = rn * T[0]
    A[N - 1, :] = 0.0; A[N - 1, :N] = rn * T[N - 1]
    A[N, :] = 0.0; A[N, N:] = rn * T[0]
    for c in (0, N - 1, N, 2 * N - 1):
        A[:, c] *= 0.5
    rhs = np.zeros(2 * N)
    rhs[N - 1] = math.sqrt(4.0 * math.pi)
    fac = 1.0 / np.abs(A).max(axis=1)
    x = np.linalg.solve(fac[:, None] * A, fac * rhs)
    C = rad["C"]
    return C @ x[:N], C @ x[N:]


_INFO_ROWS = ("s", "z", "b", "j", "wp", "p0")


class _Section:
    """Context of one kernel section: names the calls recorded into a launch
    plan, or brackets eager launches with HIP events when timers are on.  A
    module-level class, so that a section holds the run without forming a
    reference cycle (a run must be freed as soon as its last reference goes,
    not at some later garbage collection)."""

    __slots__ = ("run", "name", "prev", "ev")

    def __init__(self, run, name):
        self.run, self.name = run, name

    def __enter__(self):
        rec = _lib.recording()
        if rec is not None:
            self.prev = rec.section
            rec.section = self.name
        elif self.run.timers is not None:
            self.ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            self.ev[0].record()
        return self

    def __exit__(self, *exc):
        rec = _lib.recording()
        if rec is not None:
            rec.section = self.prev
        elif self.run.timers is not None:
            self.ev[1].record()
            self.run.timers.setdefault(self.name, []).append(self.ev)
        return False


class DynamoRun:
    """State and step of one Boussinesq Chebyshev CNAB2 dynamo run on the GPU."""

    def __init__(self, cfg, device="cuda", res: Resolution | None = None, fused_grid: bool = True):
        for _ in self._construct(cfg, device, res, fused_grid):
            pass

    @classmethod
    def staged(cls, cfg, device="cuda", res: Resolution | None = None, fused_grid: bool = True):
        """(run, stages): the run object and a generator that builds it in three
        stages (buffers, operator storage, job tables), one per next(), so that
        a server can interleave the construction with other work."""
        run = cls.__new__(cls)
        return run, run._construct(cfg, device, res, fused_grid)

    def _construct(self, cfg, device, res, fused_grid):
        self.p = p = cfg if isinstance(cfg, Params) else Params(cfg)
        p.check_step_supported()
        # anelastic runs (strat > 0) take the separate grid path (FP64 FFTs and a
        # grid kernel of their own) with FP64 transforms
        self.anel = bool(p.l_anel)
        if self.anel:
            fused_grid = False
        self.device = require_gpu(device)
        self.res = res if res is not None else Resolution(p, self.device)
        r = self.res
        L, N, M, T, P = r.L, r.N, r.M, r.T, r.P
        self.L, self.N = L, N
        self.Nb = N - 2
        dev = self.device
        # every (L, cols) complex array of the run is a view of one zeroed pool,
        # each starting on a 256-byte boundary
        slot = lambda cols: -(-(L * cols) // 16) * 16
        n_arr = len(_FIELDS) + 4 + 4 * len(_SYS) + 4 + 11
        self._pool = torch.zeros(n_arr * slot(N) + slot(2 * N), dtype=C128, device=dev)
        used = [0]

        def z(cols=N):
            a = self._pool[used[0]:used[0] + L * cols].view(L, cols)
            used[0] += slot(cols)
            return a

        self.f = {k: z() for k in _FIELDS}
        self.tmp = {k: z() for k in ("d2s", "d2z", "dddw", "ddddw")}
        self.old = {k: z() for k in _SYS}
        self.impl = {k: z() for k in _SYS}
        self.expl_cur = {k: z() for k in _SYS}
        self.expl_prev = {k: z() for k in _SYS}
        self.rhs = {k: z() for k in ("s", "z", "b", "j")}
        self.rhs["wp"] = z(2 * N)
        # radial-loop spectral outputs (boundary columns stay zero)
        self.nl = {k: z() for k in ("Advr", "Advt", "Advp", "dVSr", "VSt", "VxBr", "VxBt",
                                    "VxBp", "dVSr_dr", "dVxBt_dr", "dAdvt_dr")}
        # grid-space work (FP64) between Fourier buffers F[c][m][t][b].  Fused
        # (default): one kernel for inverse FFT + products + forward FFT, no grid in
        # HBM.  Separate: grids G[c][b][t][phi] and three kernels.
        self.fused_grid = fused_grid
        # mixed precision: Legendre transforms on FP32 MFMA with FP32 accumulation
        # (or BF16 MFMA on split operands), and (fused grid path) FP32 Fourier
        # buffers, azimuthal FFTs and nonlinear products; solves, radial
        # derivatives, IMEX terms and diagnostics stay FP64
        self.leg_prec = LEGENDRE_PREC[p.precision]
        if self.leg_prec == 1:
            r.f32_tables()   # shared FP32 tables, converted once, outside any plan
        elif self.leg_prec >= 2:
            r.bf16_tables(self.leg_prec)
        self.grid_generic = False  # force the run-time FFT plan of the fused kernel
        fdt = torch.float32 if (self.leg_prec and fused_grid) else F64
        self.F_s = r.fourier_buffer(1, N, fdt)
        self.F_v = r.fourier_buffer(3, 4 * N, fdt)
        if fused_grid:
            self.F_o = r.fourier_buffer(9, self.Nb, fdt)
        else:
            self.G_s = r.grid_buffer(1, N)
            self.G_v = r.grid_buffer(3, 4 * N)
            self.G_o = r.grid_buffer(9, self.Nb)
            nFo = 9 * M * T * self.Nb * 2
            self.F_o = self.F_v[:nFo] if nFo <= self.F_v.numel() else r.fourier_buffer(9, self.Nb)
        assert used[0] == self._pool.numel()
        if self.anel:
            self._anel_buffers()
        # finite-difference radial scheme: the double-curl poloidal equation
        # (params.py:143); operators affine in the implicit weight, assembled on
        # the host once (magic_amd.fdstep), refreshed on the device per dt change
        self.dc = bool(p.l_finite_diff)
        if self.dc:
            self._dc_setup()
        self.cfl = torch.zeros(2 * N, dtype=torch.int64, device=dev)
        self.am = None
        if p.l_correct_AMz or p.l_correct_AMe:   # update_z.py:450-475
            from . import anelastic
            prof = self.prof if self.anel else dict(
                rho0=np.ones(N), beta=np.zeros(N), dbeta=np.zeros(N))
            am = anelastic.am_setup(p, r.rad, prof)
            self.am = {k: (torch.as_tensor(v, dtype=F64).to(dev) if isinstance(v, np.ndarray)
                           else v) for k, v in am.items()}
            self.am["l1m0"] = int(r.maps["lm_start"][0] + 1)
            # (l=1, m=1) is the first row of order 1 (absent when minc > 1)
            self.am["l1m1"] = int(r.maps["lm_start"][1]) if (p.minc == 1 and p.n_m > 1) else -1
        self.cfl_host = torch.zeros(2 * N, dtype=torch.int64).pin_memory()
        # conducting inner core (sigma_ratio > 0): the inner fields on the
        # even-Chebyshev grid, their old states and implicit terms (BPR353: one
        # implicit slot per stage; the inner explicit terms of a non-rotating
        # core are zero, update_b.py:875-901), and the coupled right-hand sides
        # (L, N + n_ic) of the B and J systems
        self.cond_ic = bool(p.l_cond_ic and p.l_mag)
        if self.cond_ic:
            Nic, NT = p.n_r_ic_max, p.n_r_tot
            zi = lambda cols=Nic: torch.zeros(L, cols, dtype=C128, device=dev)
            self.fic = {k: zi() for k in ("b_ic", "db_ic", "ddb_ic", "aj_ic", "dj_ic", "ddj_ic")}
            self.old_ic = {k: zi() for k in ("b", "j")}
            self.impl_ic = {k: zi() for k in ("b", "j")}
            dirk = p.time_scheme.upper() == "BPR353"
            self.implS_ic = [self.impl_ic] + [{k: zi() for k in ("b", "j")} for _ in range(3)] \
                if dirk else [self.impl_ic]
            self.rhs_c = {k: zi(NT) for k in ("b", "j")}
            self.ic = grid.inner_core(p, r.rad["r_icb"])
            self.d_ic = {"O_r_ic": torch.as_tensor(self.ic["O_r_ic"], dtype=F64).to(dev)}
        # rotating conducting inner core (nRotIC = 1, no-slip ICB): omega_ic is
        # the ICB value of z(l=1, m=0) times c_z10_omega; its angular-momentum
        # equation is an IMEX scalar whose terms live in `rot` ([0] omega_ic,
        # [1] old state, [2..5] implicit terms, [6..9] explicit terms: the Lorentz
        # torque).  The radial loop adds, at the ICB level only, the torque (a
        # grid sum of br bphi), the induction of the rigidly rotating core (the
        # theta component of v x B, analysed into VxBt at the ICB) and the
        # advection of the inner field (the inner explicit terms, one array per
        # explicit slot of the scheme: zeros written every radial loop when the
        # core does not rotate, as the reference writes them); the z(l=1, m=0)
        # solve uses its own operator (update_z.py:244-298)
        self.rot_ic = bool(self.cond_ic and p.nRotIC == 1)
        if self.cond_ic:
            from . import inner_core
            self.rot_c = inner_core.rotation_constants(p)
            self.rot = torch.zeros(16, dtype=F64, device=dev)
            self.expl_ic = [{k: zi() for k in ("b", "j")} for _ in range(4 if dirk else 2)]
            self.lm10 = int(r.maps["lm_start"][0] + 1)
        if self.rot_ic:
            self.F_icb = r.fourier_buffer(3, 1)
            self.G_icb = r.grid_buffer(3, 1)
            self.F2_icb = r.fourier_buffer(2, 1)
            self.G2_icb = r.grid_buffer(2, 1)
            upd = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=F64).to(dev)
            self.d_ic["gauss"] = upd(r.th["gauss_grid"])
            self.d_ic["sin2"] = upd(r.th["sin_grid"] ** 2)
            self.d_ic["lm10"] = torch.tensor([self.lm10], dtype=I32, device=dev)
        # launch shapes (timed once per shape, same results whatever the choice):
        # radial levels per workgroup of the grid step, waves per workgroup of
        # the Legendre synthesis
        from . import tune
        self.grid_rb = 0
        if fused_grid:
            self.grid_rb = tune.grid_level_block(r, N, p, fdt,
                                                 bufs=(self.F_s, self.F_v, self.F_o, self.cfl))
        self.syn_vec_shape = tune.synth_shape(r, N, self.leg_prec, fdt, bufs=(self.f["s"], self.F_v))
        self.syn_scal_shape = tune.synth_shape(r, N, self.leg_prec, fdt, bufs=(self.f["s"], self.F_s),
                                               vector=False)
        self.ana_waves = tune.analysis_waves(r, N, self.leg_prec, fdt,
                                             bufs=(self.F_o, self.nl["Advr"])) if fused_grid else 0
        yield
        self._build_operator_storage()
        yield
        self._build_jobs()
        # row-operator tile shape (32 x 64 or 32 x 128), timed on the S, Z, B, J
        # solves (the fields are still zero: nothing to overwrite)
        self.rowop_wide = tune.rowop_wide(self)
        self.dt_hist = [0.0, 0.0]
        self.dtMax = 0.0
        self.wimp_lin0 = None
        self.n_rebuild = 0
        self.time = 0.0
        self.timers = None   # {name: [(start_event, end_event), ...]} while timing
        # launch plans: the radial and lm phases as recorded C-ABI calls, one per
        # parity of the explicit-history buffers (they alternate every step);
        # the dt-dependent IMEX weights come from a device buffer, so a plan (and
        # its HIP graph) stays valid for the whole run
        self._E = (self.expl_cur, self.expl_prev)
        self._par = 0
        # IMEX weights on the device: [0..3] the rhs weights of old, impl, expl0,
        # expl1, [4] the buoyancy factor wimp_lin0 * BuoFac, [8..11] the weights
        # of the further terms of a BPR353 stage
        self.wdev = torch.zeros(16, dtype=F64, device=dev)
        self.w_host = torch.zeros(16, dtype=F64).pin_memory()
        # BPR353 (time_scheme.py:157-326): four stages per step, explicit terms in
        # slots 0..2 (slot 3 is never written: stage 4's pressure rhs reads it as
        # zeros, as the reference does), implicit terms one slot ahead, the old
        # state written after stage 4; weights per stage in wdev_stage
        self.dirk = p.time_scheme.upper() == "BPR353"
        self.stage = 0
        if self.dirk:
            zl = lambda: {k: torch.zeros(L, N, dtype=C128, device=dev) for k in _SYS}
            self.implS = [self.impl, zl(), zl(), zl()]
            self.explS = [self.expl_cur, self.expl_prev, zl(), zl()]
            self.wdev_stage = torch.zeros(4 * 16, dtype=F64, device=dev)
            self.w_stage_host = torch.zeros(4 * 16, dtype=F64).pin_memory()
        self._plans = {}
        self._graphs = {}
        self._cap_stream = None
        self.use_graphs = False
        self._sec_ev = None
        self._cfl_event = None
        self.dt_log = []         # dt used by each step
        self.dt_schedule = None  # optional step sizes to use instead of the Courant decision
        self.n_step = 0          # steps taken
        self._ctl = None         # device step control buffers (run_steps / step_async)
        self.expl_pending = False  # a radial loop's explicit terms await the lm loop

    def _dc_setup(self):
        """Profiles of the double-curl kernels on the device and the derivative
        operator of the flux d/dr (orho1 r^2 Advt) (boundary values dropped)."""
        from . import fdstep
        rad, N, dev = self.res.rad, self.N, self.device
        self.dc_prof = self.prof if self.anel else fdstep.boussinesq_profile(rad)
        pr = self.dc_prof
        up = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=F64).to(dev)
        self.d_dc = {k: up(pr[k]) for k in ("beta", "dbeta", "ddbeta", "orho1", "rho0", "visc",
                                             "dLvisc", "ddLvisc", "rgrav")}
        wt = pr["orho1"] * rad["r"] ** 2
        wt[0] = wt[N - 1] = 0.0
        Op = rad["D1"] * wt[None, :]
        Rp = -(-N // 64) * 64
        K4 = -(-N // 4) * 4
        mt = np.zeros((K4, Rp))
        mt[:N, :N] = Op.T
        self.MT_dvxvh = (up(mt), N, Rp)

    def _dc_args(self):
        a = _lib.DcArgs()
        a.lc, a.rc = self._lmcoef(), self._radcoef()
        d, f, t = self.d_dc, self.f, self.tmp
        if self.anel:
            for k in ("beta", "dbeta", "ddbeta", "orho1", "rho0", "visc", "dLvisc", "ddLvisc"):
                setattr(a, k, ptr(d[k]))
        a.grav = ptr(d["rgrav"])
        a.w, a.dw, a.ddw, a.dddw, a.ddddw = (ptr(f["w"]), ptr(f["dw"]), ptr(f["ddw"]),
                                             ptr(t["dddw"]), ptr(t["ddddw"]))
        a.z, a.dz, a.s = ptr(f["z"]), ptr(f["dz"]), ptr(f["s"])
        a.CorFac, a.BuoFac, a.lm_l0 = self.p.CorFac, self.p.BuoFac, 0
        return a

    def _dc_storage(self):
        """Host-assembled operators A0 + w A1 (S, Z (B, J) stacked like mat_nn;
        the double-curl W in the first (l_max+1) N^2 of mat_wp) and the fixed
        l = 0 pressure block of the W row operator."""
        from . import fdstep
        p, rad, N, dev = self.p, self.res.rad, self.N, self.device
        lp1 = p.l_max + 1
        ops = fdstep.band_operators(p, rad, self.dc_prof)
        up = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=F64).to(dev)
        (a0, a1, ks), (w0, w1, kw) = ops["szbj"], ops["w"]
        # banded operators (LINPACK storage): A0 + w A1, factorised in place, then
        # solved against the identity for the dense inverses the folds need
        self.dc_ops = dict(a0=up(a0.reshape(-1)), a1=up(a1.reshape(-1)), w0=up(w0.reshape(-1)),
                           w1=up(w1.reshape(-1)), ks=ks, kw=kw)
        self.dc_abd = torch.empty(a0.size, dtype=F64, device=dev)
        self.dc_abd_w = torch.empty(w0.size, dtype=F64, device=dev)
        self.dc_piv = torch.zeros(self.nsys * lp1, N, dtype=I32, device=dev)
        self.dc_piv_w = torch.zeros(lp1, N, dtype=I32, device=dev)
        eye = torch.zeros(lp1 * N, N, dtype=C128, device=dev)
        eye.view(lp1, N, N).diagonal(dim1=1, dim2=2).fill_(1.0)
        self.dc_eye = eye
        self.dc_x = torch.empty(lp1 * N, N, dtype=C128, device=dev)
        self.dc_l = torch.as_tensor(np.repeat(np.arange(lp1), N), dtype=I32).to(dev)
        D1, D2 = rad["D1"], rad["D2"]
        self.Op4dc = up(np.concatenate([np.eye(N), D1, D2, D1 @ D2]))   # (4N, N)
        self.Op22 = up(D2 @ D2)
        # MT[l][k][r] of degree 0: p(l=0) = inv_p0 @ (pressure half of the rhs)
        mt0 = np.zeros((self.K4wp, self.Rpad6))
        mt0[N:2 * N, 4 * N:5 * N] = fdstep.p0_inverse(p, rad, self.dc_prof).T
        self.MT["wp"][0].copy_(up(mt0))

    def _rebuild_launches_dc(self, wimp_lin0, ctl=None):
        """Finite-difference rebuild: A0 + w A1 for every system in band storage,
        the batched LINPACK band LU of all degrees (the reference's banded
        path, ``update_s.py:147-163``, ``update_w_doublecurl.py:224-242``), band
        solves against the identity for the dense inverses, and the folds with
        the identity and D1, D2 (S, Z, B, J) and with I, D1, D2, D1 D2 and D2 D2
        (W) into the row operators; rows 1 and N-2 of the W input (its boundary
        conditions on w') are dropped, as the reference zeroes those rhs rows."""
        N, lp1, ns = self.N, self.p.l_max + 1, self.nsys
        st = stream_handle()
        wdev, when = (None, None) if ctl is None else ctl
        info = lambda key: ptr(self.gj_info[_INFO_ROWS.index(key)])
        o, ks, kw = self.dc_ops, self.dc_ops["ks"], self.dc_ops["kw"]
        X, nx = ptr(self.dc_x), lp1 * N
        band = lambda abd, piv, k, nbatch, inf: _lib.call(
            "mh_band_factor_gated", abd, piv, inf, N, k, k, nbatch, when, st)
        inverse = lambda abd, piv, k: _lib.call(   # X rows l N + c: column c of inv_l
            "mh_band_solve_gated", abd, piv, None, None, ptr(self.dc_l), None, ptr(self.dc_eye),
            X, nx, N, k, k, when, st)
        _lib.call("mh_affine_f64", ptr(o["a0"]), ptr(o["a1"]), ptr(self.dc_abd),
                  self.dc_abd.numel(), wimp_lin0, wdev, when, st)
        band(ptr(self.dc_abd), ptr(self.dc_piv), ks, ns * lp1, info("s"))
        per = (3 * ks + 1) * N
        sb = lp1 * self.K4 * self.Rpad3
        for sy in range(ns):
            inverse(ptr(self.dc_abd) + 8 * sy * lp1 * per, ptr(self.dc_piv) + 4 * sy * lp1 * N, ks)
            # MT[l][c][r] = sum_j Op3[r][j] inv_l[j][c], inv_l[j][c] = Re X[l N + c][j]
            self._gemm(X, 2 * N, 2, 2 * N * N, ptr(self.Op3), 1, N, 0,
                       ptr(self.MT4) + 8 * sy * sb, self.Rpad3, 1, self.K4 * self.Rpad3,
                       N, 3 * N, N, lp1, when)
        self._zero(self.MT4[1:ns, 0], when, count=ns - 1, stride=sb)
        # W, degrees 1..l_max (degree 0 holds the fixed l = 0 pressure block)
        nb = lp1 - 1
        _lib.call("mh_affine_f64", ptr(o["w0"]), ptr(o["w1"]), ptr(self.dc_abd_w),
                  self.dc_abd_w.numel(), wimp_lin0, wdev, when, st)
        band(ptr(self.dc_abd_w), ptr(self.dc_piv_w), kw, lp1, info("wp"))
        inverse(ptr(self.dc_abd_w), ptr(self.dc_piv_w), kw)
        MT = self.MT["wp"]
        sbw = self.K4wp * self.Rpad6
        mt1 = ptr(MT[1])
        x1 = X + 8 * 2 * N * N     # degree 1
        self._gemm(x1, 2 * N, 2, 2 * N * N, ptr(self.Op4dc), 1, N, 0, mt1, self.Rpad6, 1, sbw,
                   N, 4 * N, N, nb, when)
        self._gemm(x1, 2 * N, 2, 2 * N * N, ptr(self.Op22), 1, N, 0, mt1 + 8 * 5 * N, self.Rpad6,
                   1, sbw, N, N, N, nb, when)
        for k in (1, N - 2):
            self._zero(MT[1:, k], when, count=nb, stride=sbw)

    def _anel_buffers(self):
        """Fourier and grid buffers of the anelastic radial loop, the reference
        state's profiles on the device, and the angular-momentum and l = 0
        pressure factors."""
        from . import anelastic
        r, p, N, L, dev = self.res, self.p, self.N, self.L, self.device
        Nb = N - 2
        self.prof = prof = anelastic.background(p, r.rad)
        up = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=F64).to(dev)
        self.d_an = {k: up(prof[k]) for k in ("beta", "dbeta", "rho0", "orho1", "otemp1",
                                              "dLtemp0", "visc", "dLvisc", "kappa", "dLkappa")}
        self.d_an["rgrav_rho"] = up(prof["rho0"] * prof["rgrav"])
        self.d_an["cosn2"] = up(r.th["cos_grid"] * r.th["ost2_grid"])
        self.d_an["p0u"] = up(anelastic.p0_weights(p, r.rad, prof))
        self.p0_integral = anelastic.p0_integral(p, prof)
        self.F_s2, self.G_s2 = r.fourier_buffer(1, 2 * N), r.grid_buffer(1, 2 * N)
        self.F_va, self.G_va = r.fourier_buffer(3, 3 * N), r.grid_buffer(3, 3 * N)
        if p.l_mag:
            self.F_vb, self.G_vb = r.fourier_buffer(3, 2 * N), r.grid_buffer(3, 2 * N)
        self.F_d, self.G_d = r.fourier_buffer(2, N), r.grid_buffer(2, N)
        self.F_oa, self.G_oa = r.fourier_buffer(10, Nb), r.grid_buffer(10, Nb)
        self.heatLM = torch.zeros(L, N, dtype=C128, device=dev)
        self.zeroLN = torch.zeros(L, N, dtype=C128, device=dev)
        self.comp_scale_10 = torch.tensor([0, 0, 0, 0, 1, 1, 1, 1, 1, 1], dtype=I32, device=dev)

    def _anel_prof(self, a):
        """Fill an MhAnelProf from the run's device profiles."""
        d = self.d_an
        for k in ("beta", "dbeta", "rho0", "orho1", "otemp1", "dLtemp0", "visc", "dLvisc",
                  "kappa", "dLkappa", "rgrav_rho"):
            setattr(a, k, ptr(d[k]))
        a.ViscHeatFac, a.OhmLossFac = self.prof["ViscHeatFac"], self.prof["OhmLossFac"]

    # ------------------------------------------------------------ timing
    def enable_timers(self, on: bool = True):
        """Record HIP events around each kernel section (on the launch stream)."""
        self.timers = {} if on else None

    def _t(self, name):
        return _Section(self, name)

    def _on_section(self, name, begin):
        if begin:
            self._sec_ev = torch.cuda.Event(enable_timing=True)
            self._sec_ev.record()
        else:
            end = torch.cuda.Event(enable_timing=True)
            end.record()
            self.timers.setdefault(name, []).append((self._sec_ev, end))

    def _plan(self, kind):
        """The recorded launch plan of a phase for the current parity (for
        BPR353: for the current stage)."""
        if kind == "step":   # a whole step: one plan per parity (CNAB2) or the four stages
            key = (kind, "dirk") if self.dirk else (kind, self._par)
        else:
            key = (kind, "stage", self.stage) if self.dirk else (kind, self._par)
        plan = self._plans.get(key)
        if plan is None:
            with _lib.record() as rec:
                rec.plan = True   # a step's launch plan (replayed every step)
                if kind == "radial":
                    self._radial_launches()
                elif kind == "lm":
                    self._lm_launches()
                else:
                    self._step_launches()
            plan = self._plans[key] = rec.calls
        return key, plan

    def _graph(self, key, plan):
        """The captured HIP graph of a recorded plan (captured once).

        Captured on the current stream when it is not the device's default
        stream (a run served on its own stream), else on a capture stream of
        this run; no device-wide synchronisation (torch.cuda.graph would stall
        every other run on the GPU).  The garbage collector is off during the
        capture: nothing may free device memory or destroy graphs while a
        stream capture is open."""
        g = self._graphs.get(key)
        if g is None:
            cur = torch.cuda.current_stream()
            cap = cur
            if cur == torch.cuda.default_stream(self.device):
                if self._cap_stream is None:
                    self._cap_stream = torch.cuda.Stream(device=self.device)
                cap = self._cap_stream
                cap.wait_stream(cur)
            g = torch.cuda.CUDAGraph()
            gc_on = gc.isenabled()
            gc.disable()
            try:
                with torch.cuda.stream(cap):
                    g.capture_begin()
                    try:
                        _lib.run_plan(plan, stream_handle())
                    finally:
                        g.capture_end()
            finally:
                if gc_on:
                    gc.enable()
            if cap is not cur:
                cur.wait_stream(cap)
            self._graphs[key] = g
        return g

    def close(self):
        """Release the run now: wait for its queued work, destroy its graphs and
        drop its device buffers (a server retiring a run, or a benchmark
        between cases).  The run cannot step afterwards."""
        torch.cuda.current_stream().synchronize()
        self._graphs.clear()
        self._plans.clear()
        for k in ("_pool", "f", "tmp", "old", "impl", "expl_cur", "expl_prev", "rhs", "nl", "_E",
                  "F_s", "F_v", "F_o", "MT", "mat_nn", "mat_wp", "gj_work", "gj_scal", "jobs1",
                  "jobs2", "jobs_td", "jobs_setup", "_ctl"):
            if hasattr(self, k):
                setattr(self, k, None)
        self.closed = True

    def prepare_graph(self, kind, par):
        """Record and capture the launch plan of phase `kind` ("radial", "lm" or
        "step") for explicit-history parity `par`, without running it."""
        saved = (self._par, self.expl_cur, self.expl_prev)
        self._par, self.expl_cur, self.expl_prev = par, self._E[par], self._E[1 - par]
        try:
            key, plan = self._plan(kind)
            if self.use_graphs:
                self._graph(key, plan)
        finally:
            self._par, self.expl_cur, self.expl_prev = saved

    def _exec(self, kind):
        """Run the radial or lm phase: recorded once per parity, then replayed
        eagerly (with section timers when enabled) or as a captured HIP graph."""
        key, plan = self._plan(kind)
        if self.use_graphs and self.timers is None:
            self._graph(key, plan).replay()
        else:
            _lib.run_plan(plan, stream_handle(),
                          self._on_section if self.timers is not None else None)

    def timer_summary(self):
        """{name: (launch count, mean ms)} of the recorded sections (synchronises)."""
        torch.cuda.synchronize()
        out = {}
        for k, evs in (self.timers or {}).items():
            ms = [a.elapsed_time(b) for a, b in evs]
            out[k] = (len(ms), float(np.mean(ms)))
        return out

    # ------------------------------------------------------------ operators
    def _build_operator_storage(self):
        r, N, dev = self.res, self.N, self.device
        lp1 = self.p.l_max + 1
        self.K4 = -(-N // 4) * 4
        self.K4wp = -(-(2 * N) // 4) * 4
        self.Rpad3 = -(-(3 * N) // 64) * 64
        self.Rpad6 = -(-(6 * N) // 64) * 64
        # S, Z, B, J operators stacked [system][l][k][r] so that one rebuild launch
        # covers the four systems
        self.nsys = 4 if self.p.l_mag else 2
        self.MT4 = torch.zeros(4, lp1, self.K4, self.Rpad3, dtype=F64, device=dev)
        self.MT = {k: self.MT4[i] for i, k in enumerate(("s", "z", "b", "j"))}
        self.MT["wp"] = torch.zeros(lp1, self.K4wp, self.Rpad6, dtype=F64, device=dev)
        ops = self._radial_operators()
        self.Op3, self.Opw, self.Opp, self.cheb = ops["Op3"], ops["Opw"], ops["Opp"], ops["cheb"]
        if self.cond_ic:
            Nic, NT = self.p.n_r_ic_max, self.p.n_r_tot
            self.K4c = -(-NT // 4) * 4
            rp = lambda R: -(-R // 64) * 64
            self.MTc = {k: {part: (torch.zeros(lp1, self.K4c, rp(3 * n), dtype=F64, device=dev),
                                   3 * n, rp(3 * n))
                            for part, n in (("oc", N), ("ic", Nic))} for k in ("b", "j")}
            k4i = -(-Nic // 4) * 4
            mt = np.zeros((k4i, rp(2 * Nic)))
            mt[:Nic, :2 * Nic] = np.concatenate([self.ic["D1"], self.ic["D2"]]).T
            self.MT_ic12 = (torch.as_tensor(mt, dtype=F64).to(dev), 2 * Nic, rp(2 * Nic))
        if self.rot_ic:
            self.MT_z10 = torch.zeros(self.K4, self.Rpad3, dtype=F64, device=dev)
        if self.cond_ic:
            # coupled B, J (and z(l=1, m=0)) matrices as A0 + w A1, refreshed and
            # inverted on the device at each rebuild
            from . import inner_core
            up = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=F64).to(dev)
            NT = self.p.n_r_tot
            (b0, b1), (j0, j1) = inner_core.coupled_affine(self.p, r.rad, self.ic)
            self.ic_ops = dict(a0=up(np.stack([b0, j0]).reshape(-1)),
                               a1=up(np.stack([b1, j1]).reshape(-1)),
                               op_ic=up(inner_core.inner_fold(self.ic)))
            self.mat_c = torch.empty(2 * lp1, NT, NT, dtype=F64, device=dev)
            self.gj_work_c = torch.empty(2 * lp1 * 2 * NT * NT if NT > 208 else 1, dtype=F64,
                                         device=dev)
            self.gj_scal_c = torch.empty(2 * lp1 * 2 * NT, dtype=F64, device=dev)
            if self.rot_ic:
                z0, z1 = inner_core.z10_affine(self.p, r.rad, self.rot_c["c_dt_z10"])
                self.ic_ops.update(z0=up(z0), z1=up(z1))
                self.mat_z10 = torch.empty(N, N, dtype=F64, device=dev)
        self.MT_d1, self.MT_d1r2, self.MT_d12, self.MT_d123 = (
            ops[k] for k in ("MT_d1", "MT_d1r2", "MT_d12", "MT_d123"))
        # dense collocation stacks of the Chebyshev rebuild (the finite-difference
        # one factorises band storage instead)
        nd = 1 if self.dc else lp1
        self.mat_nn = torch.empty(self.nsys * nd, N, N, dtype=F64, device=dev)
        self.mat_wp = torch.empty(nd, 2 * N, 2 * N, dtype=F64, device=dev)
        self.mat_p0 = torch.empty(N, N, dtype=F64, device=dev)
        # Gauss-Jordan work space: only the kernel for n > 208 uses it
        nwork = 0
        if 2 * N > 208:
            nwork = nd * 2 * (2 * N) ** 2
        if N > 208:
            nwork = max(nwork, self.nsys * nd * 2 * N * N)
        self.gj_work = torch.empty(max(nwork, 1), dtype=F64, device=dev)
        self.gj_scal = torch.empty(max(self.nsys * lp1, 2 * lp1) * 2 * N * 2, dtype=F64, device=dev)
        # first zero pivot per degree, one row per system (s, z, b, j, wp, p0)
        self.gj_info = torch.zeros(len(_INFO_ROWS), lp1, dtype=I32, device=dev)
        self._info_host = torch.zeros(len(_INFO_ROWS), lp1, dtype=I32).pin_memory()
        self._info_event = None
        self.defer_checks = False  # check the pivots of a rebuild at the next step (no host wait)
        if self.dc:
            self._dc_storage()
        elif self.anel:
            # anelastic Chebyshev operators: S, Z and (w, p) affine in the implicit
            # weight (assembled once on the host), the l = 0 pressure block fixed
            from . import anelastic
            ops = anelastic.affine_operators(self.p, self.res.rad, self.prof)
            up = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=F64).to(dev)
            (a0, a1), (w0, w1) = ops["sz"], ops["wp"]
            self.anel_ops = dict(a0=up(a0.reshape(-1)), a1=up(a1.reshape(-1)),
                                 w0=up(w0.reshape(-1)), w1=up(w1.reshape(-1)),
                                 mt0=up(anelastic.p0_block(self.p, self.res.rad, self.prof,
                                                           self.K4wp, self.Rpad6)))
    def _radial_operators(self):
        """Radial operators of the resolution, uploaded once and shared by every
        run of it (read-only): the DCT-I / derivative stacks folded into the
        inverses, the Chebyshev collocation matrices of the builders, and the
        uniform radial-derivative operators (transposed, padded)."""
        r, N, dev = self.res, self.N, self.device
        ops = getattr(r, "_radial_ops", None)
        if ops is not None:
            return ops
        rad = r.rad
        CT, D1, D2, D3 = rad["CT"], rad["D1"], rad["D2"], rad["D3"]
        up = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=F64).to(dev)
        ops = dict(Op3=up(np.concatenate([CT, D1 @ CT, D2 @ CT])),            # (3N, N)
                   Opw=up(np.concatenate([CT, D1 @ CT, D2 @ CT, D3 @ CT])),   # (4N, N)
                   Opp=up(np.concatenate([CT, D1 @ CT])),                     # (2N, N)
                   cheb={k: up(rad[k]) for k in ("T", "T1", "T2", "T3")})
        r2 = rad["r2"]

        def mt(blocks):  # list of (N, N) blocks stacked as output rows -> MT[k][r]
            Op = np.concatenate(blocks)
            R = Op.shape[0]
            Rp = -(-R // 64) * 64
            out = np.zeros((self.K4, Rp))
            out[:N, :R] = Op.T
            return up(out), R, Rp

        ops.update(MT_d1=mt([D1]), MT_d1r2=mt([D1 * r2[None, :]]), MT_d12=mt([D1, D2]),
                   MT_d123=mt([D1, D2, D3]))
        r._radial_ops = ops
        return ops

    def _gemm(self, A, sAr, sAc, sAb, B, sBr, sBc, sBb, Cp, sCr, sCc, sCb, M, Nn, K, batch,
              when=None):
        g = _lib.GemmArgs()
        g.when = when
        g.A, g.B, g.C = A, B, Cp
        g.sAr, g.sAc, g.sAb, g.sBr, g.sBc, g.sBb, g.sCr, g.sCc, g.sCb = (
            sAr, sAc, sAb, sBr, sBc, sBb, sCr, sCc, sCb)
        g.M, g.Nn, g.K, g.batch, g.alpha = M, Nn, K, batch, 1.0
        _lib.call("mh_gemm_f64", _lib.ref(g), stream_handle())

    def _build_all(self, wl0, ctl=None):
        """Collocation matrices of every system in one launch: S, Z (B, J)
        stacked in mat_nn, (w,p) in mat_wp, the l=0 pressure system in mat_p0."""
        p, rad = self.p, self.res.rad
        a = _lib.BuildArgs()
        a.R, a.D1, a.D2, a.D3 = (ptr(self.cheb[k]) for k in ("T", "T1", "T2", "T3"))
        a.or1, a.or2 = ptr(self.res.d["or1"]), ptr(self.res.d["or2"])
        a.out, a.out_wp, a.out_p0 = ptr(self.mat_nn), ptr(self.mat_wp), ptr(self.mat_p0)
        a.N, a.l_max, a.n_cheb, a.nsys = self.N, p.l_max, p.n_cheb_max, self.nsys
        a.ktopv, a.kbotv, a.kind = p.ktopv, p.kbotv, 6
        a.rnorm, a.wimp, a.opr, a.opm = rad["rnorm"], wl0, p.opr, p.opm
        if ctl is not None:
            a.wimp_dev, a.when = ctl
        _lib.call("mh_build_matrices", _lib.ref(a), stream_handle())

    def _zero(self, t, when=None, count=1, stride=0):
        n = t[0].numel() if count > 1 else t.numel()
        _lib.call("mh_fill_f64", ptr(t), n, count, stride, 0.0, when, stream_handle())

    def build_all_matrices(self, wimp_lin0: float):
        """Rebuild every per-degree operator for the implicit weight ``wimp_lin0``.

        Device-side replacement of ``build_all_matrices`` (reference
        ``step_time.py:1530-1548``): assemble, equilibrate and invert the
        collocation matrices of all degrees at once, then fold the DCT-I and
        the radial derivatives into the inverses.
        """
        self._rebuild_launches(wimp_lin0)
        self.wimp_lin0 = wimp_lin0
        self.n_rebuild += 1
        self._queue_info_check()

    def _coupled_launches(self, wimp_lin0, ctl=None):
        """The coupled outer/inner B and J row operators of a conducting inner core
        (and the z(l=1, m=0) operator of a rotating one) for ``wimp_lin0``:
        A0 + w A1, batched inverses with row scaling (degree 0 dropped), and the
        folds with the outer (I, D1, D2) and inner (I, D1, D2) operators."""
        p, N, lp1, st = self.p, self.N, self.p.l_max + 1, stream_handle()
        Nic, NT = p.n_r_ic_max, p.n_r_tot
        wdev, when = (None, None) if ctl is None else ctl
        o = self.ic_ops
        nn = NT * NT
        _lib.call("mh_affine_f64", ptr(o["a0"]), ptr(o["a1"]), ptr(self.mat_c), 2 * lp1 * nn,
                  wimp_lin0, wdev, when, st)
        _lib.call("mh_gj_inverse", ptr(self.mat_c), ptr(self.mat_c), ptr(self.gj_work_c),
                  ptr(self.gj_scal_c), NT, 2 * lp1, 1,
                  ptr(self.gj_info[_INFO_ROWS.index("b")]), when, st)
        for kk, key in enumerate(("b", "j")):
            A = ptr(self.mat_c) + 8 * kk * lp1 * nn
            for part, op, n_in, off in (("oc", self.Op3, N, 0), ("ic", o["op_ic"], Nic, N)):
                MT, R, Rp = self.MTc[key][part]
                # MT[l][k][r] = sum_j op[r][j] inv_l[off + j][k]
                self._gemm(A + 8 * off * NT, 1, NT, nn, ptr(op), 1, n_in, 0, ptr(MT), Rp, 1,
                           self.K4c * Rp, NT, R, n_in, lp1, when)
                self._zero(MT[0], when)
        if self.rot_ic:
            _lib.call("mh_affine_f64", ptr(o["z0"]), ptr(o["z1"]), ptr(self.mat_z10), N * N,
                      wimp_lin0, wdev, when, st)
            _lib.call("mh_gj_inverse", ptr(self.mat_z10), ptr(self.mat_z10), ptr(self.gj_work),
                      ptr(self.gj_scal), N, 1, 1,
                      ptr(self.gj_info[_INFO_ROWS.index("p0")]) + 4, when, st)
            self._gemm(ptr(self.mat_z10), 1, N, 0, ptr(self.Op3), 1, N, 0, ptr(self.MT_z10),
                       self.Rpad3, 1, 0, N, 3 * N, N, 1, when)

    def _rebuild_launches(self, wimp_lin0, ctl=None):
        """The launches of a rebuild (ten for the Boussinesq Chebyshev systems;
        the anelastic, finite-difference and coupled inner-core operators add
        their own).  ctl = (wimp_dev, when) device pointers: the weight is read on the device
        and every kernel returns at once unless *when != 0 (the rebuild inside
        a captured step)."""
        if self.dc:
            self._rebuild_launches_dc(wimp_lin0, ctl)
        else:
            self._rebuild_launches_cheb(wimp_lin0, ctl)
        if self.cond_ic:   # after the others: its pivot flags fill the B and J rows
            self._coupled_launches(wimp_lin0, ctl)

    def _rebuild_launches_cheb(self, wimp_lin0, ctl=None):
        N, lp1, ns = self.N, self.p.l_max + 1, self.nsys
        st = stream_handle()
        when = None if ctl is None else ctl[1]
        info = lambda key: ptr(self.gj_info[_INFO_ROWS.index(key)])
        self._build_all(wimp_lin0, ctl)
        if self.anel:   # the anelastic S, Z and (w, p) matrices replace the Boussinesq ones
            o, wdev = self.anel_ops, (None if ctl is None else ctl[0])
            _lib.call("mh_affine_f64", ptr(o["a0"]), ptr(o["a1"]), ptr(self.mat_nn),
                      2 * lp1 * N * N, wimp_lin0, wdev, when, st)
            _lib.call("mh_affine_f64", ptr(o["w0"]), ptr(o["w1"]), ptr(self.mat_wp),
                      lp1 * 4 * N * N, wimp_lin0, wdev, when, st)
        # S, Z (B, J): all degrees of all systems in one inverse and one fold;
        # their pivot-flag rows are consecutive
        _lib.call("mh_gj_inverse", ptr(self.mat_nn), ptr(self.mat_nn), ptr(self.gj_work),
                  ptr(self.gj_scal), N, ns * lp1, 1, info("s"), when, st)
        # MT[k][l][c][r] = sum_j inv_kl[j][c] Op3[r][j]
        self._gemm(ptr(self.mat_nn), 1, N, N * N, ptr(self.Op3), 1, N, 0,
                   ptr(self.MT4), self.Rpad3, 1, self.K4 * self.Rpad3, N, 3 * N, N, ns * lp1, when)
        # Z, B, J have no l = 0 equation
        sb = lp1 * self.K4 * self.Rpad3
        self._zero(self.MT4[1:ns, 0], when, count=ns - 1, stride=sb)
        # coupled (w, p)
        _lib.call("mh_gj_inverse", ptr(self.mat_wp), ptr(self.mat_wp), ptr(self.gj_work),
                  ptr(self.gj_scal), 2 * N, lp1, 2, info("wp"), when, st)
        MT = self.MT["wp"]
        sb = self.K4wp * self.Rpad6
        self._gemm(ptr(self.mat_wp), 1, 2 * N, 4 * N * N, ptr(self.Opw), 1, N, 0,
                   ptr(MT), self.Rpad6, 1, sb, 2 * N, 4 * N, N, lp1, when)
        self._gemm(ptr(self.mat_wp) + 8 * N * 2 * N, 1, 2 * N, 4 * N * N, ptr(self.Opp), 1, N, 0,
                   ptr(MT) + 8 * 4 * N, self.Rpad6, 1, sb, 2 * N, 2 * N, N, lp1, when)
        self._zero(MT[0], when)
        # l = 0 pressure: p(l=0) from the p0 system acting on the pressure half
        p0 = self.mat_p0
        _lib.call("mh_gj_inverse", ptr(p0), ptr(p0), ptr(self.gj_work), ptr(self.gj_scal), N, 1, 1,
                  info("p0"), when, st)
        self._gemm(ptr(p0), 1, N, 0, ptr(self.Opp), 1, N, 0,
                   ptr(MT) + 8 * (N * self.Rpad6 + 4 * N), self.Rpad6, 1, 0, N, 2 * N, N, 1, when)
        if self.anel:   # the anelastic l = 0 pressure block (host inverse, fixed)
            o = self.anel_ops
            _lib.call("mh_affine_f64", ptr(o["mt0"]), ptr(o["mt0"]), ptr(MT[0]),
                      self.K4wp * self.Rpad6, 0.0, None, when, st)

    def _queue_info_check(self):
        """Queue the host copy of the pivot flags of all systems; they are checked
        now or (defer_checks) at the next step."""
        self._info_host.copy_(self.gj_info, non_blocking=True)
        if self._info_event is None:
            self._info_event = torch.cuda.Event()
        self._info_event.record()
        self._info_pending = True
        if not self.defer_checks:
            self._check_info()

    def _check_info(self):
        """Raise the reference's AssertionError for the first singular system
        (``update_s.py:152`` message form) of the last rebuild."""
        if not getattr(self, "_info_pending", False):
            return
        self._info_event.synchronize()
        self._info_pending = False
        info = self._info_host.numpy()
        lp1 = self.p.l_max + 1
        for row, name in enumerate(_INFO_ROWS):
            if name in ("b", "j") and not self.p.l_mag:
                continue
            n = 1 if name == "p0" else lp1
            bad = np.nonzero(info[row, :n])[0]
            if bad.size:
                l = int(bad[0])
                raise AssertionError(f"Singular {name} matrix for l={l}, info={int(info[row, l])}")

    # ------------------------------------------------------------ job tables
    def _build_jobs(self):
        """Row-operator job descriptors and tile lists (pointers are fixed per run)."""
        r, N, L, lp1 = self.res, self.N, self.L, self.p.l_max + 1
        rows = r.d["l_rows"]
        roff = r.maps["l_rows_off"]

        def job(X, ldx, rows_ptr, nrows, MT, K, R, Rpad, Y, ldy, rseg):
            """Job columns; rows_ptr, nrows and MT may be per-job arrays."""
            return dict(X=X, ldx=ldx, rows=rows_ptr, nrows=nrows, MT=MT, K=K, R=R, Rpad=Rpad,
                        Y=list(Y), ldy=ldy, rseg=rseg)

        def upload(specs):
            """(job records, {tile columns: (tile list, count)}) on the device."""
            out = {}
            for tc in (64, 128):
                arr, tiles = pack_row_jobs(specs, tc)
                out[tc] = (torch.as_tensor(tiles.reshape(-1)).to(self.device), int(tiles.shape[0]))
            dj = torch.frombuffer(bytearray(arr.tobytes()), dtype=torch.uint8).to(self.device)
            return dj, out

        f, tmp = self.f, self.tmp
        # lm loop, stage 1: S, Z, B, J per degree
        jobs = []
        systems = [("s", self.rhs["s"], (f["s"], f["ds"], tmp["d2s"])),
                   ("z", self.rhs["z"], (f["z"], f["dz"], tmp["d2z"]))]
        if self.p.l_mag and not self.cond_ic:
            systems += [("b", self.rhs["b"], (f["b"], f["db"], f["ddb"])),
                        ("j", self.rhs["j"], (f["aj"], f["dj"], f["ddj"]))]
        ls = np.arange(lp1, dtype=np.int64)
        row_ptrs = ptr(rows) + 4 * np.asarray(roff[:lp1], np.int64)
        nrows_l = np.diff(np.asarray(roff, np.int64))   # orders of each degree
        for key, X, Y in systems:
            MT = self.MT[key]
            rp_k, nr_k = row_ptrs, nrows_l
            if key == "z" and self.rot_ic:
                # (l=1, m=0), the first row of degree 1, has its own operator
                rp_k, nr_k = row_ptrs.copy(), nrows_l.copy()
                rp_k[1] += 4
                nr_k[1] -= 1
                jobs.append(job(ptr(X), N, ptr(self.d_ic["lm10"]), 1, ptr(self.MT_z10), N, 3 * N,
                                self.Rpad3, [ptr(y) for y in Y], N, N))
            jobs.append(job(ptr(X), N, rp_k, nr_k, ptr(MT) + 8 * ls * self.K4 * self.Rpad3,
                            N, 3 * N, self.Rpad3, [ptr(y) for y in Y], N, N))
        if self.p.l_mag and self.cond_ic:
            # coupled B and J: outer (f, f', f'') and inner (f, f', f'') outputs of
            # one (N + n_ic)-long right-hand side per mode
            Nic, NT = self.p.n_r_ic_max, self.p.n_r_tot
            fi = self.fic
            for key, Yo, Yi in (("b", (f["b"], f["db"], f["ddb"]),
                                 (fi["b_ic"], fi["db_ic"], fi["ddb_ic"])),
                                ("j", (f["aj"], f["dj"], f["ddj"]),
                                 (fi["aj_ic"], fi["dj_ic"], fi["ddj_ic"]))):
                for part, Y, ld in (("oc", Yo, N), ("ic", Yi, Nic)):
                    MT, R, Rp = self.MTc[key][part]
                    jobs.append(job(ptr(self.rhs_c[key]), NT, row_ptrs, nrows_l,
                                    ptr(MT) + 8 * ls * self.K4c * Rp, NT, R, Rp,
                                    [ptr(y) for y in Y], ld, ld))
        self.jobs1 = upload(jobs)
        # lm loop, stage 2: coupled (w, p) per degree; double curl: w and its four
        # derivatives from the w half, the l = 0 pressure from the pressure half
        Y = [f["w"], f["dw"], f["ddw"], tmp["dddw"], f["p"], tmp["ddddw"] if self.dc else f["dp"]]
        self.jobs2 = upload([job(ptr(self.rhs["wp"]), 2 * N, row_ptrs, nrows_l,
                                 ptr(self.MT["wp"]) + 8 * ls * self.K4wp * self.Rpad6, 2 * N,
                                 6 * N, self.Rpad6, [ptr(y) for y in Y], N, N)])
        # radial loop: derivatives of the flux terms
        nl = self.nl
        jobs = [job(ptr(nl["dVSr"]), N, 0, L, ptr(self.MT_d1[0]), N, N, self.MT_d1[2],
                    [ptr(nl["dVSr_dr"])], N, N)]
        if self.p.l_mag:
            jobs.append(job(ptr(nl["VxBt"]), N, 0, L, ptr(self.MT_d1r2[0]), N, N,
                            self.MT_d1r2[2], [ptr(nl["dVxBt_dr"])], N, N))
        if self.dc:   # d/dr (orho1 r^2 Advt) of the double-curl explicit term
            jobs.append(job(ptr(nl["Advt"]), N, 0, L, ptr(self.MT_dvxvh[0]), N, N,
                            self.MT_dvxvh[2], [ptr(nl["dAdvt_dr"])], N, N))
        self.jobs_td = upload(jobs)
        # set-up: derivatives of the initial fields
        jobs = [job(ptr(f["s"]), N, 0, L, ptr(self.MT_d12[0]), N, 2 * N, self.MT_d12[2],
                    [ptr(f["ds"]), ptr(tmp["d2s"])], N, N),
                job(ptr(f["z"]), N, 0, L, ptr(self.MT_d12[0]), N, 2 * N, self.MT_d12[2],
                    [ptr(f["dz"]), ptr(tmp["d2z"])], N, N),
                job(ptr(f["w"]), N, 0, L, ptr(self.MT_d123[0]), N, 3 * N, self.MT_d123[2],
                    [ptr(f["dw"]), ptr(f["ddw"]), ptr(tmp["dddw"])], N, N)]
        if self.dc:   # the set-up leaves dp as it is (step_time.py:465-470)
            self.jobs_setup2 = upload([job(ptr(f["ddw"]), N, 0, L, ptr(self.MT_d12[0]), N,
                                           2 * N, self.MT_d12[2],
                                           [ptr(tmp["dddw"]), ptr(tmp["ddddw"])], N, N)])
        else:
            jobs.append(job(ptr(f["p"]), N, 0, L, ptr(self.MT_d1[0]), N, N, self.MT_d1[2],
                            [ptr(f["dp"])], N, N))
        if self.p.l_mag:
            jobs += [job(ptr(f["b"]), N, 0, L, ptr(self.MT_d12[0]), N, 2 * N, self.MT_d12[2],
                         [ptr(f["db"]), ptr(f["ddb"])], N, N),
                     job(ptr(f["aj"]), N, 0, L, ptr(self.MT_d12[0]), N, 2 * N, self.MT_d12[2],
                         [ptr(f["dj"]), ptr(f["ddj"])], N, N)]
        if self.p.l_mag and self.cond_ic:
            Nic, fi, (MT, R, Rp) = self.p.n_r_ic_max, self.fic, self.MT_ic12
            jobs += [job(ptr(fi[k]), Nic, 0, L, ptr(MT), Nic, R, Rp,
                         [ptr(fi[d1]), ptr(fi[d2])], Nic, Nic)
                     for k, d1, d2 in (("b_ic", "db_ic", "ddb_ic"), ("aj_ic", "dj_ic", "ddj_ic"))]
        self.jobs_setup = upload(jobs)

    def _rowop(self, jt, wide=None):
        dj, tiles = jt
        wide = self.rowop_wide if wide is None else wide
        tl, nt = tiles[128 if wide else 64]
        _lib.call("mh_rowop", ptr(dj), ptr(tl), nt, int(wide), stream_handle())

    # ------------------------------------------------------------ coefficient structs
    def _lmcoef(self):
        d = self.res.d
        c = _lib.LmCoef()
        c.l, c.m, c.dLh = ptr(d["l"]), ptr(d["m"]), ptr(d["dLh"])
        c.dTh2A, c.dTh2S, c.dTh3A, c.dTh3S = (ptr(d[k]) for k in ("dTh2A", "dTh2S", "dTh3A", "dTh3S"))
        c.dTh4A, c.dTh4S = ptr(d["dTh4A"]), ptr(d["dTh4S"])
        c.l_max, c.L = self.p.l_max, self.L
        return c

    def _radcoef(self):
        d = self.res.d
        c = _lib.RadCoef()
        c.or1, c.or2, c.or3, c.or4 = (ptr(d[k]) for k in ("or1", "or2", "or3", "or4"))
        c.r2, c.rgrav, c.N = ptr(d["r2"]), ptr(d["rgrav"]), self.N
        return c

    # ------------------------------------------------------------ init
    def initialize_fields(self, noise_amp: float | None = None, noise_seed: int | None = None):
        """``initialize_fields`` (reference ``init_fields.py:470-532``) + optional noise."""
        p, r, f = self.p, self.res, self.f
        for t in f.values():
            t.zero_()
        up = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=C128).to(self.device)
        rr, r_cmb, r_icb = r.rad["r"], r.rad["r_cmb"], r.rad["r_icb"]
        def lm(l, m):
            if m % p.minc or m > l or l > p.l_max:
                raise ValueError(f"initial mode (l={l}, m={m}) is not in the truncation "
                                 f"(l_max={p.l_max}, minc={p.minc})")
            return int(r.maps["lm_start"][m // p.minc] + (l - m))
        # the fields are zero but for a few rows and the rank-one seeded noise,
        # so only those go to the device (no (L, N) host arrays): s = rows + noise
        amp = p.noise_amp if noise_amp is None else noise_amp
        seed = p.noise_seed if noise_seed is None else noise_seed
        if amp:
            coef, shape = noise_factors(p, r, seed, amp)
            f["s"].add_(up(coef)[:, None] * up(shape)[None, :])
        if p.l_heat:
            if self.dc:
                from . import fdstep
                s0, p0 = fdstep.conduction_state(p, r.rad, self.dc_prof)
            elif self.anel:
                from . import anelastic
                s0, p0 = anelastic.conduction_state(p, r.rad, self.prof)
            else:
                s0, p0 = conduction_state(p, r)
            f["s"][0].add_(up(s0))
            f["p"][0].copy_(up(p0))
            if p.init_s1 >= 100:
                l, m = p.init_s1 // 100, p.init_s1 % 100
                x = 2.0 * rr - r_cmb - r_icb
                f["s"][lm(l, m)].add_(up(p.amp_s1 * (1.0 - 3.0 * x ** 2 + 3.0 * x ** 4 - x ** 6)))
        if p.l_mag and self.cond_ic:
            # initB with a conducting inner core (init_fields.py:426-456)
            for t in self.fic.values():
                t.zero_()
            r_ic = self.ic["r_ic"]
            b_pol = p.amp_b1 * math.sqrt(3.0 * math.pi) / (3.0 + r_cmb)
            b_tor = -4.0 / 3.0 * p.amp_b1 * math.sqrt(math.pi / 5.0)
            f["b"][lm(1, 0)].copy_(up(b_pol * (rr ** 3 - 4.0 / 3.0 * r_cmb * rr ** 2)))
            self.fic["b_ic"][lm(1, 0)].copy_(up(b_pol * r_icb ** 2 * (
                0.5 * r_ic ** 2 / r_icb + 0.5 * r_icb - 4.0 / 3.0 * r_cmb)))
            f["aj"][lm(2, 0)].copy_(up(b_tor * rr * np.sin(math.pi * (rr / r_cmb))))
            arg = math.pi * r_icb / r_cmb
            sa, ca = math.sin(arg), math.cos(arg)
            aj1 = (arg - 2.0 * sa * ca) / (arg + sa * ca)
            aj2 = (1.0 - aj1) * r_icb * sa / ca
            self.fic["aj_ic"][lm(2, 0)].copy_(up(b_tor * (
                aj1 * r_ic * np.sin(math.pi * r_ic / r_cmb) + aj2 * np.cos(math.pi * r_ic / r_cmb))))
        elif p.l_mag:
            b_pol = p.amp_b1 * math.sqrt(3.0 * math.pi) / 4.0
            b_tor = -4.0 / 3.0 * p.amp_b1 * math.sqrt(math.pi / 5.0)
            f["b"][lm(1, 0)].copy_(up(b_pol * (rr ** 3 - 4.0 / 3.0 * r_cmb * rr ** 2
                                               + r_icb ** 4 / (3.0 * rr))))
            f["aj"][lm(2, 0)].copy_(up(b_tor * rr * np.sin(math.pi * (rr - r_icb))))
        if self.rot_ic and p.omega_ic1 != 0.0:
            # initV (init_fields.py:528-532): the core starts rotating at
            # omega_ic1, z(l=1, m=0) at the ICB set to match
            f["z"][lm(1, 0), self.N - 1] = p.omega_ic1 / self.rot_c["c_z10_omega"]

    def setup_initial_state(self):
        """Derivatives, old states and implicit terms from the fields (``step_time.py:398-617``)."""
        self._par = 0 if self.expl_cur is self._E[0] else 1
        extra = (self.explS[2:] + self.implS[1:]) if self.dirk else []
        for d in [self.old, self.impl, self.expl_cur, self.expl_prev] + extra:
            for t in d.values():
                t.zero_()
        self._rowop(self.jobs_setup)
        if self.dc:
            self._rowop(self.jobs_setup2)
        self._impl(zero_edges=True)
        # the reference's set-up fills the magnetic implicit terms on the
        # interior levels only (step_time.py:582-595)
        for k in ("b", "j"):
            self.impl[k][:, 0] = 0.0
            self.impl[k][:, -1] = 0.0
        if self.cond_ic:   # get_mag_ic_rhs_imp (update_b.py:803-873)
            for d in [self.old_ic] + self.implS_ic:
                for t in d.values():
                    t.zero_()
            self._ic_impl(self.old_ic, self.impl_ic)
        if self.cond_ic:
            for d in self.expl_ic:
                for t in d.values():
                    t.zero_()
            self.rot.zero_()
        if self.rot_ic:   # step_time.py:597-617
            self._ic_omega_update(True, 2)

    def _ic_impl(self, old, impl):
        """Inner-core old states (old: None for none) and implicit terms."""
        p, fi = self.p, self.fic
        a = _lib.IcImplArgs()
        a.b, a.db, a.ddb = ptr(fi["b_ic"]), ptr(fi["db_ic"]), ptr(fi["ddb_ic"])
        a.aj, a.dj, a.ddj = ptr(fi["aj_ic"]), ptr(fi["dj_ic"]), ptr(fi["ddj_ic"])
        a.old_b, a.old_j = (None, None) if old is None else (ptr(old["b"]), ptr(old["j"]))
        a.impl_b, a.impl_j = ptr(impl["b"]), ptr(impl["j"])
        a.dLh, a.l, a.O_r_ic = ptr(self.res.d["dLh"]), ptr(self.res.d["l"]), ptr(self.d_ic["O_r_ic"])
        a.L, a.n_ic = self.L, p.n_r_ic_max
        a.fac, a.or2_icb = p.opm * p.O_sr, float(self.res.rad["or2"][-1])
        _lib.call("mh_ic_impl", _lib.ref(a), stream_handle())

    def dt_cap(self, dt: float) -> float:
        """dt clamped to intfac * E for rotating runs (``step_time.py:1564-1570``)."""
        if not self.p.ek < 0.0:
            dt = min(dt, self.p.int_fac * self.p.ek)
        return dt

    def initialize_dt(self, dt: float):
        """dt <= intfac * E for rotating runs (``step_time.py:1554-1571``)."""
        dt = self.dt_cap(dt)
        self.dt_hist = [dt, dt]
        self.dtMax = dt

    # ------------------------------------------------------------ kernels
    def _impl(self, old=None, impl=None, zero_edges=False):
        """Implicit terms and old states of the new fields into `impl` / `old`
        (default: the run's own arrays)."""
        old = self.old if old is None else old
        impl = self.impl if impl is None else impl
        a = _lib.ImplArgs()
        a.lc, a.rc = self._lmcoef(), self._radcoef()
        f, t = self.f, self.tmp
        for k, v in (("s", f["s"]), ("ds", f["ds"]), ("d2s", t["d2s"]), ("z", f["z"]),
                     ("d2z", t["d2z"]), ("w", f["w"]), ("dw", f["dw"]), ("ddw", f["ddw"]),
                     ("dddw", t["dddw"]), ("p", f["p"]), ("dp", f["dp"]), ("b", f["b"]),
                     ("ddb", f["ddb"]), ("aj", f["aj"]), ("ddj", f["ddj"])):
            setattr(a, k, ptr(v))
        for k in _SYS:
            setattr(a, "old_" + k, ptr(old[k]))
            setattr(a, "impl_" + k, ptr(impl[k]))
        a.opr, a.opm, a.BuoFac = self.p.opr, self.p.opm, self.p.BuoFac
        a.l_mag, a.lm_l0 = int(self.p.l_mag), 0
        if self.anel:
            # the set-up zeroes the w, p terms at l = 0 and the boundaries, the lm
            # loop leaves them (step_time.py:528-535, update_wp.py:500-531)
            A = _lib.ImplAnelArgs()
            A.base = a
            self._anel_prof(A.prof)
            A.dz = ptr(f["dz"])
            A.zero_edges = int(zero_edges)
            _lib.call("mh_impl_anel", _lib.ref(A), stream_handle())
        else:
            _lib.call("mh_impl", _lib.ref(a), stream_handle())
        if self.dc:   # w from the double-curl operator; no pressure terms
            d = self._dc_args()
            d.old_w, d.impl_w = ptr(old["w"]), ptr(impl["w"])
            d.old_p, d.impl_p = ptr(old["p"]), ptr(impl["p"])
            _lib.call("mh_impl_dc", _lib.ref(d), stream_handle())

    def radial_loop(self):
        """Explicit terms of the current state; returns (dtrkc_min, dthkc_min)."""
        self._exec("radial")
        self.expl_pending = True
        return self._cfl_decision()

    def _radial_launches(self):
        if self.anel:
            return self._radial_launches_anel()
        r, f, N, Nb = self.res, self.f, self.N, self.Nb
        # 1. Legendre synthesis
        with self._t("leg_syn_scal"):
            r.synth([[(f["s"], None, 0)]], N, N, self.F_s, N, vector=False, prec=self.leg_prec,
                    waves=self.syn_scal_shape[0], nct=self.syn_scal_shape[1])
        groups = [[(f["b"], None, 1), (f["db"], None, 0), (f["aj"], None, 0)],
                  [(f["aj"], None, 1), (f["dj"], None, 0), (f["ddb"], f["b"], 2)],
                  [(f["w"], None, 1), (f["dw"], None, 0), (f["z"], None, 0)],
                  [(f["z"], None, 1), (f["dz"], None, 0), (f["ddw"], f["w"], 2)]]
        with self._t("leg_syn_vec"):
            r.synth(groups, N, N, self.F_v, 4 * N, vector=True, prec=self.leg_prec,
                    waves=self.syn_vec_shape[0], nct=self.syn_vec_shape[1])
        d = r.d
        if self.fused_grid:
            # 2-4. fused inverse FFT + nonlinear products + Courant + forward FFT
            with self._t("grid_step"):
                r.grid_step(self.F_s, self.F_v, self.F_o, N, self.cfl, self.p,
                            generic=self.grid_generic, rb=self.grid_rb)
        else:
            self._grid_separate()
        # 5. Legendre analysis into (L, N) arrays at columns 1..N-2
        nl = self.nl
        off = 16  # one complex column
        with self._t("leg_ana_scal"):
            r.analysis(self.F_o, Nb, Nb, [0, 1, 2],
                       [ptr(nl["Advr"]) + off, ptr(nl["dVSr"]) + off, ptr(nl["VxBr"]) + off],
                       [None, None, None], N, vector=False, prec=self.leg_prec,
                       waves=self.ana_waves)
        with self._t("leg_ana_vec"):
            r.analysis(self.F_o, Nb, Nb, [(3, 4), (5, 6), (7, 8)],
                       [ptr(nl["Advt"]) + off, ptr(nl["VSt"]) + off, ptr(nl["VxBt"]) + off],
                       [ptr(nl["Advp"]) + off, None, ptr(nl["VxBp"]) + off], N, vector=True,
                       prec=self.leg_prec, waves=self.ana_waves)
        if self.cond_ic:
            with self._t("ic_expl"):
                self._icb_launches()
        self._td_launches()

    def _rot_args(self):
        """MhIcRotArgs with the run's rotation state and constants."""
        p, rad, c = self.p, self.res.rad, self.rot_c
        a = _lib.IcRotArgs()
        a.rot, a.lm10, a.L, a.N, a.n_ic = ptr(self.rot), self.lm10, self.L, self.N, p.n_r_ic_max
        a.T, a.P = self.res.T, self.res.P
        a.c_z10_omega, a.c_dt_z10 = c["c_z10_omega"], c["c_dt_z10"]
        a.c_lorentz, a.fac_lt = c["c_lorentz"], c["fac_lt"]
        a.or1_icb, a.or2_icb = float(rad["or1"][-1]), float(rad["or2"][-1])
        a.or4_icb, a.r_icb = float(rad["or4"][-1]), float(rad["r"][-1])
        for q in range(8):
            a.term[q] = -1
        a.expl_slot, a.impl_slot = 6, 2
        return a

    def _ic_expl_slot(self):
        """Explicit slot the radial loop writes: the parity's newest (CNAB2) or
        stage - 1 (BPR353)."""
        return max(0, self.stage - 1) if self.dirk else self._par

    def _icb_launches(self):
        """The inner core's part of the radial loop (step_time.py:1126-1138,
        1181-1192; update_b.py:875-901).  Rotating core: (br, btheta, bphi) on the
        ICB grid (FP64 synthesis and inverse FFT of the ICB level), the Lorentz
        torque and the grid of the rotation's induction, its spheroidal analysis
        into VxBt at the ICB.  Every conducting core: the inner explicit terms
        (the advection of the inner field; zeros without rotation)."""
        r, f, N = self.res, self.f, self.N
        a = self._rot_args()
        k = self._ic_expl_slot()
        a.expl_slot = 6 + k
        if self.rot_ic:
            icb = lambda t: t[:, N - 1:]
            r.synth([[(icb(f["b"]), None, 1), (icb(f["db"]), None, 0), (icb(f["aj"]), None, 0)]],
                    1, N, self.F_icb, 1, vector=True, prec=0)
            r.fft(False, self.F_icb, 1, self.G_icb, 3, 1)
            n = r.T * r.P
            a.br, a.bp = ptr(self.G_icb), ptr(self.G_icb) + 8 * 2 * n
            a.vt, a.vp = ptr(self.G2_icb), ptr(self.G2_icb) + 8 * n
            a.gauss, a.sin2 = ptr(self.d_ic["gauss"]), ptr(self.d_ic["sin2"])
            _lib.call("mh_ic_torque", _lib.ref(a), stream_handle())
            r.fft(True, self.F2_icb, 1, self.G2_icb, 2, 1, comp_scale=r.d["comp_scale_2"])
            r.analysis(self.F2_icb, 1, 1, [(0, 1)], [ptr(self.nl["VxBt"]) + 16 * (N - 1)], [None],
                       N, vector=True, prec=0)
        e = self.expl_ic[k]
        a.b_ic, a.aj_ic = ptr(self.fic["b_ic"]), ptr(self.fic["aj_ic"])
        a.e_b, a.e_j = ptr(e["b"]), ptr(e["j"])
        a.m, a.dLh = ptr(r.d["m"]), ptr(r.d["dLh"])
        _lib.call("mh_ic_advect", _lib.ref(a), stream_handle())

    def _ic_omega_rhs(self, wdev):
        """omega's IMEX right-hand side into the ICB entry of the z(l=1, m=0)
        right-hand side (update_z.py:355-371)."""
        a = self._rot_args()
        a.wdev, a.rhs_z = wdev, ptr(self.rhs["z"])
        if self.dirk:
            # old, I0, E0, E1, I1, E2, I2, I3 as the stage's weights (rhs layout)
            k = self.stage
            terms = [1, 2, 6] + ([7, 3] if k >= 2 else [-1, -1]) + \
                    ([8, 4] if k >= 3 else [-1, -1]) + ([5] if k >= 4 else [-1])
        else:
            terms = [1, 2, 6 + self._par, 7 - self._par, -1, -1, -1, -1]
        for q, t in enumerate(terms):
            a.term[q] = t
        _lib.call("mh_ic_omega_rhs", _lib.ref(a), stream_handle())

    def _ic_omega_update(self, store_old, impl_slot):
        """omega_ic, the old state (store_old) and the implicit term of omega's
        equation from the solved z(l=1, m=0) (update_z.py:435-501)."""
        a = self._rot_args()
        a.store_old, a.impl_slot = int(store_old), impl_slot
        a.z, a.dz = ptr(self.f["z"]), ptr(self.f["dz"])
        _lib.call("mh_ic_omega_update", _lib.ref(a), stream_handle())

    @property
    def omega_ic(self) -> float:
        """Rotation rate of the inner core (0.0 without a rotating core; reads
        the device state: synchronises the run's stream)."""
        return float(self.rot[0].item()) if self.cond_ic else 0.0

    def _radial_launches_anel(self):
        """The anelastic radial loop (``radial_loop_anel``, step_time.py:1215-1475),
        FP64: synthesis of s and l(l+1) z; of u, du/dr and the gradient of the
        radial velocity; of B and curl B; the azimuthal derivatives of u_theta,
        u_phi in Fourier space; inverse FFTs; the anelastic grid products with the
        Courant maxima; forward FFTs and analysis of the bulk levels; the explicit
        terms, with the heating and 1/rho0 in the entropy equation.  Mixed
        precision changes the Legendre transforms only (FP64 Fourier buffers,
        FFTs and products)."""
        r, f, N, Nb, p = self.res, self.f, self.N, self.Nb, self.p
        M, T = r.M, r.T
        with self._t("leg_syn_scal"):
            r.synth([[(f["s"], None, 0)], [(f["z"], None, 1)]], N, N, self.F_s2, 2 * N,
                    vector=False, prec=self.leg_prec)
        groups = [[(f["w"], None, 1), (f["dw"], None, 0), (f["z"], None, 0)],
                  [(f["dw"], None, 1), (f["ddw"], None, 0), (f["dz"], None, 0)],
                  [(self.zeroLN, None, 0), (f["w"], None, 1), (self.zeroLN, None, 0)]]
        with self._t("leg_syn_vec"):
            r.synth(groups, N, N, self.F_va, 3 * N, vector=True, prec=self.leg_prec)
            if p.l_mag:
                r.synth([[(f["b"], None, 1), (f["db"], None, 0), (f["aj"], None, 0)],
                         [(f["aj"], None, 1), (f["dj"], None, 0), (f["ddb"], f["b"], 2)]],
                        N, N, self.F_vb, 2 * N, vector=True, prec=self.leg_prec)
        d = _lib.DphiArgs()
        d.F, (d.Fc_stride, d.Fm_stride), d.FB, d.c0 = ptr(self.F_va), r.fstrides(3 * N), 3 * N, 1
        d.out, (d.Oc_stride, d.Om_stride), d.OB = ptr(self.F_d), r.fstrides(N), N
        d.B, d.n_m, d.T, d.minc = N, M, T, p.minc
        with self._t("fft_inv"):
            _lib.call("mh_fourier_dphi", _lib.ref(d), stream_handle())
            r.fft(False, self.F_s2, 2 * N, self.G_s2, 1, 2 * N)
            r.fft(False, self.F_va, 3 * N, self.G_va, 3, 3 * N)
            if p.l_mag:
                r.fft(False, self.F_vb, 2 * N, self.G_vb, 3, 2 * N)
            r.fft(False, self.F_d, N, self.G_d, 2, N)
        a = _lib.NlAnelArgs()
        npt = r.T * r.P
        a.Gs, a.Gv, a.Gv_c_stride = ptr(self.G_s2), ptr(self.G_va), 3 * N * npt
        if p.l_mag:
            a.Gb, a.Gb_c_stride = ptr(self.G_vb), 2 * N * npt
        a.Gd, a.Gd_c_stride = ptr(self.G_d), N * npt
        a.Go, a.Go_c_stride = ptr(self.G_oa), Nb * npt
        a.N, a.T, a.P, a.l_mag = N, r.T, r.P, int(p.l_mag)
        dd = r.d
        a.or1, a.or2, a.or4 = ptr(dd["or1"]), ptr(dd["or2"]), ptr(dd["or4"])
        a.ost2, a.cosn2 = ptr(dd["ost2_grid"]), ptr(self.d_an["cosn2"])
        a.valri2, a.valhi2 = ptr(dd["valri2"]), ptr(dd["valhi2"])
        self._anel_prof(a.prof)
        a.LFfac, a.cf2, a.af2 = p.LFfac, p.cour_fac ** 2, p.alf_fac ** 2
        a.alf_damp, a.cfl = int(p.l_cour_alf_damp), ptr(self.cfl)
        with self._t("get_nl"):
            _lib.call("mh_nl_anel", _lib.ref(a), stream_handle())
        with self._t("fft_fwd"):
            r.fft(True, self.F_oa, Nb, self.G_oa, 10, Nb, comp_scale=self.comp_scale_10)
        nl = self.nl
        off = 16  # one complex column
        with self._t("leg_ana_scal"):
            r.analysis(self.F_oa, Nb, Nb, [0, 1, 2],
                       [ptr(nl["Advr"]) + off, ptr(nl["dVSr"]) + off, ptr(nl["VxBr"]) + off],
                       [None, None, None], N, vector=False, prec=self.leg_prec)
            r.analysis(self.F_oa, Nb, Nb, [3], [ptr(self.heatLM) + off], [None], N,
                       vector=False, prec=self.leg_prec)
        with self._t("leg_ana_vec"):
            r.analysis(self.F_oa, Nb, Nb, [(4, 5), (6, 7), (8, 9)],
                       [ptr(nl["Advt"]) + off, ptr(nl["VSt"]) + off, ptr(nl["VxBt"]) + off],
                       [ptr(nl["Advp"]) + off, None, ptr(nl["VxBp"]) + off], N, vector=True,
                       prec=self.leg_prec)
        self._td_launches()
        e = _lib.EntropyAnelArgs()
        e.es, e.heat, e.orho1 = ptr(self.expl_cur["s"]), ptr(self.heatLM), ptr(self.d_an["orho1"])
        e.L, e.N = self.L, N
        with self._t("td"):
            _lib.call("mh_entropy_anel", _lib.ref(e), stream_handle())

    def _grid_separate(self):
        r, N, Nb, d = self.res, self.N, self.Nb, self.res.d
        # 2. inverse FFT
        with self._t("fft_inv"):
            r.fft(False, self.F_s, N, self.G_s, 1, N)
            r.fft(False, self.F_v, 4 * N, self.G_v, 3, 4 * N)
        # 3. nonlinear products + Courant maxima
        a = _lib.NlArgs()
        a.Gs, a.Gv, a.Gv_c_stride = ptr(self.G_s), ptr(self.G_v), 4 * N * r.T * r.P
        a.Go, a.Go_c_stride = ptr(self.G_o), Nb * r.T * r.P
        a.N, a.T, a.P = N, r.T, r.P
        a.or2, a.or4, a.ost2 = ptr(d["or2"]), ptr(d["or4"]), ptr(d["ost2_grid"])
        a.valri2, a.valhi2 = ptr(d["valri2"]), ptr(d["valhi2"])
        a.LFfac, a.cf2, a.af2 = self.p.LFfac, self.p.cour_fac ** 2, self.p.alf_fac ** 2
        a.l_mag, a.alf_damp = int(self.p.l_mag), int(self.p.l_cour_alf_damp)
        a.cfl = ptr(self.cfl)
        with self._t("get_nl"):
            _lib.call("mh_get_nl", _lib.ref(a), stream_handle())
        # 4. forward FFT of the bulk products
        with self._t("fft_fwd"):
            r.fft(True, self.F_o, Nb, self.G_o, 9, Nb, comp_scale=d["comp_scale_v"])

    def _td_launches(self):
        r, f, N, nl = self.res, self.f, self.N, self.nl
        # 6. radial derivatives of the fluxes
        with self._t("rowop_dr"):
            self._rowop(self.jobs_td)
        # 7. explicit terms
        t = _lib.TdArgs()
        t.lc, t.rc = self._lmcoef(), self._radcoef()
        for k in ("Advr", "Advt", "Advp", "VSt", "VxBr", "VxBp", "dVSr_dr", "dVxBt_dr"):
            setattr(t, k, ptr(nl[k]))
        t.w, t.dw, t.z = ptr(f["w"]), ptr(f["dw"]), ptr(f["z"])
        e = self.expl_cur
        t.ew, t.ez, t.ep = ptr(e["w"]), ptr(e["z"]), ptr(e["p"])
        t.es, t.eb, t.ej = ptr(e["s"]), ptr(e["b"]), ptr(e["j"])
        t.CorFac, t.l_mag = self.p.CorFac, int(self.p.l_mag)
        with self._t("td"):
            _lib.call("mh_td", _lib.ref(t), stream_handle())
            if self.dc:
                a = self._dc_args()
                a.Advr, a.dAdvt_dr = ptr(nl["Advr"]), ptr(nl["dAdvt_dr"])
                a.ew, a.ep = ptr(e["w"]), ptr(e["p"])
                _lib.call("mh_td_dc", _lib.ref(a), stream_handle())

    def _cfl_decision(self):
        """Courant limits (host decision, like the reference's .item())."""
        self._cfl_copy()
        return self._cfl_values()

    def _cfl_copy(self):
        """Queue the copy of the per-level Courant maxima to pinned host memory."""
        self.cfl_host.copy_(self.cfl, non_blocking=True)
        if self._cfl_event is None:
            self._cfl_event = torch.cuda.Event()
        self._cfl_event.record()

    def cfl_ready(self) -> bool:
        """True once the queued Courant maxima have reached the host."""
        return self._cfl_event is None or self._cfl_event.query()

    def _cfl_values(self):
        r, N = self.res, self.N
        self._cfl_event.synchronize()
        v = self.cfl_host.numpy().view(np.float64)
        rad = r.rad
        with np.errstate(divide="ignore"):
            dtrkc = np.where(v[:N] > 0, np.sqrt(rad["delxr2"] / np.where(v[:N] > 0, v[:N], 1.0)), 1e10)
            dthkc = np.where(v[N:] > 0, np.sqrt(rad["delxh2"] / np.where(v[N:] > 0, v[N:], 1.0)), 1e10)
        return float(dtrkc.min()), float(dthkc.min())

    def lm_loop(self):
        """Implicit solves and post-processing (``_fused_lm_loop_fast`` + insulating B)."""
        p = self.p
        dt1, dt2 = self.dt_hist
        w = self.w_host
        w[0] = 1.0                                    # wimp
        w[1] = (1.0 - p.alpha) * dt1                  # wimp_lin[1]
        w[2] = (1.0 + 0.5 * dt1 / dt2) * dt1          # wexp[0]
        w[3] = -0.5 * dt1 * dt1 / dt2                 # wexp[1]
        w[4] = p.alpha * dt1 * p.BuoFac               # wimp_lin[0] * BuoFac (buoyancy)
        # the host writes w_host only after the step's Courant synchronisation, so
        # the previous step's copy has completed
        self.wdev.copy_(w, non_blocking=True)
        self._exec("lm")
        self.expl_cur, self.expl_prev = self.expl_prev, self.expl_cur
        self._par ^= 1
        self.expl_pending = False

    def _lm_launches(self):
        p = self.p
        a = _lib.RhsArgs()
        wdev = ptr(self.wdev)
        ew0 = self.expl_cur["w"]
        if self.dirk:
            # stage k: old + sum_{j<k} bexp[k][j] expl_j + sum_{j<k} bimp[k][j] impl_j
            # (time_scheme.py:240-262; weights per stage in wdev_stage)
            k = self.stage
            E, I = self.explS, self.implS
            for i, key in enumerate(_SYS):
                sy = a.sys[i]
                sy.old, sy.impl, sy.expl0 = ptr(self.old[key]), ptr(I[0][key]), ptr(E[0][key])
                if k >= 2:
                    sy.expl1, sy.t4 = ptr(E[1][key]), ptr(I[1][key])
                if k >= 3:
                    sy.t5, sy.t6 = ptr(E[2][key]), ptr(I[2][key])
                if k >= 4:
                    sy.t7 = ptr(I[3][key])
            wdev = ptr(self.wdev_stage) + 8 * 16 * (k - 1)
            ew0 = E[max(0, k - 1)]["w"]   # stage 4 reads slot 3: zeros (step_time.py:284-287)
        else:
            for i, k in enumerate(_SYS):
                a.sys[i].old, a.sys[i].impl = ptr(self.old[k]), ptr(self.impl[k])
                a.sys[i].expl0, a.sys[i].expl1 = ptr(self.expl_cur[k]), ptr(self.expl_prev[k])
        a.rs, a.rz, a.rb, a.rj, a.rwp = (ptr(self.rhs[k]) for k in ("s", "z", "b", "j", "wp"))
        a.m, a.L, a.N = ptr(self.res.d["m"]), self.L, self.N
        a.wdev = wdev
        a.bots_l0, a.l_mag = math.sqrt(4.0 * math.pi), int(p.l_mag)
        with self._t("rhs"):
            _lib.call("mh_rhs", _lib.ref(a), stream_handle())
            if self.cond_ic:
                self._ic_rhs(wdev)
            if self.rot_ic:
                self._ic_omega_rhs(wdev)
        with self._t("rowop_szbj"):
            self._rowop(self.jobs1)
        if self.am is not None:
            self._am_correct()
        bq = _lib.BuoyArgs()
        bq.rwp, bq.s, bq.ew0 = ptr(self.rhs["wp"]), ptr(self.f["s"]), ptr(ew0)
        rg = self.d_an["rgrav_rho"] if self.anel else self.res.d["rgrav"]   # rho0 rgrav
        bq.rgrav, bq.m, bq.L, bq.N = ptr(rg), ptr(self.res.d["m"]), self.L, self.N
        bq.BuoFac, bq.lm_l0, bq.wdev = p.BuoFac, 0, wdev
        with self._t("buoy"):
            if self.dc:
                a = self._dc_args()
                a.rwp, a.ew0, a.wdev = ptr(self.rhs["wp"]), ptr(ew0), wdev
                _lib.call("mh_buoy_dc", _lib.ref(a), stream_handle())
            else:
                _lib.call("mh_buoy", _lib.ref(bq), stream_handle())
            if self.anel and self.p0_integral and not self.dc:
                q0 = _lib.P0IntArgs()
                q0.rwp, q0.s, q0.u = ptr(self.rhs["wp"]), ptr(self.f["s"]), ptr(self.d_an["p0u"])
                q0.lm_l0, q0.N = 0, self.N
                _lib.call("mh_p0_integral", _lib.ref(q0), stream_handle())
        with self._t("rowop_wp"):
            self._rowop(self.jobs2)
        with self._t("impl"):
            if self.dirk:
                # implicit terms one slot ahead; the old state only after the last
                # stage (time_scheme.py:211-228): earlier stages write it into the
                # right-hand-side arrays, which the solves have consumed
                k = self.stage
                r = self.rhs
                junk = {"s": r["s"], "z": r["z"], "b": r["b"], "j": r["j"],
                        "w": r["wp"][:, :self.N], "p": r["wp"][:, self.N:]}
                self._impl(old=self.old if k == 4 else junk, impl=self.implS[k % 4])
                if self.cond_ic:
                    self._ic_impl(self.old_ic if k == 4 else None, self.implS_ic[k % 4])
                if self.rot_ic:
                    self._ic_omega_update(k == 4, 2 + k % 4)
            else:
                self._impl()
                if self.cond_ic:
                    self._ic_impl(self.old_ic, self.impl_ic)
                if self.rot_ic:
                    self._ic_omega_update(True, 2)

    def _am_correct(self):
        """Angular-momentum corrections of z(l=1, m=0) and z(l=1, m=1) and their
        derivatives after the toroidal solve (update_z.py:450-475)."""
        am, p = self.am, self.p
        a = _lib.AmArgs()
        a.z, a.dz, a.d2z = ptr(self.f["z"]), ptr(self.f["dz"]), ptr(self.tmp["d2z"])
        a.r2q, a.pz, a.pdz, a.pd2z = (ptr(am[k]) for k in ("r2q", "pz", "pdz", "pd2z"))
        a.l1m0, a.l1m1, a.N = am["l1m0"], am["l1m1"], self.N
        a.do_z, a.do_e = int(p.l_correct_AMz), int(p.l_correct_AMe)
        a.cz, a.ce = am["cz"], am["ce"]
        with self._t("am"):
            _lib.call("mh_am_correct", _lib.ref(a), stream_handle())

    def _ic_rhs(self, wdev):
        """Coupled right-hand sides of the B and J systems (mh_ic_rhs): the inner
        IMEX terms of the scheme (CNAB2 old, impl and the two explicit slots;
        BPR353 stage k the implicit slots 0..k-1 and the explicit slots 0..k-2,
        with the stage weights), the outer ones from mh_rhs."""
        a = _lib.IcRhsArgs()
        for i, key in enumerate(("b", "j")):
            sy = a.sys[i]
            sy.old = ptr(self.old_ic[key])
            if self.dirk:
                k, I = self.stage, self.implS_ic
                sy.impl = ptr(I[0][key])
                if k >= 2:
                    sy.t4 = ptr(I[1][key])
                if k >= 3:
                    sy.t6 = ptr(I[2][key])
                if k >= 4:
                    sy.t7 = ptr(I[3][key])
                E = self.expl_ic
                sy.expl0 = ptr(E[0][key])
                if k >= 2:
                    sy.expl1 = ptr(E[1][key])
                if k >= 3:
                    sy.t5 = ptr(E[2][key])
            else:
                sy.impl = ptr(self.impl_ic[key])
                sy.expl0 = ptr(self.expl_ic[self._par][key])
                sy.expl1 = ptr(self.expl_ic[1 - self._par][key])
        a.rb, a.rj = ptr(self.rhs["b"]), ptr(self.rhs["j"])
        a.out_b, a.out_j = ptr(self.rhs_c["b"]), ptr(self.rhs_c["j"])
        a.m, a.L, a.N, a.n_ic = ptr(self.res.d["m"]), self.L, self.N, self.p.n_r_ic_max
        a.wdev = wdev
        _lib.call("mh_ic_rhs", _lib.ref(a), stream_handle())

    def one_step(self, n_time_step: int, dt: float) -> float:
        """One CNAB2 step (reference ``step_time.py:1574-1637``), or one BPR353
        step (``step_time.py:1640-1680``); returns the dt used."""
        if self.dirk:
            return self._one_step_dirk(n_time_step, dt)
        self.begin_step()
        return self.finish_step(n_time_step, dt)

    # BPR353 tableaux (time_scheme.py:186-200), unscaled; row k is stage k
    BIMP = ((0.0, 0.0, 0.0, 0.0, 0.0), (1 / 2, 1 / 2, 0.0, 0.0, 0.0),
            (5 / 18, -1 / 9, 1 / 2, 0.0, 0.0), (1 / 2, 0.0, 0.0, 1 / 2, 0.0),
            (1 / 4, 0.0, 3 / 4, -1 / 2, 1 / 2))
    BEXP = ((0.0, 0.0, 0.0, 0.0, 0.0), (1.0, 0.0, 0.0, 0.0, 0.0), (4 / 9, 2 / 9, 0.0, 0.0, 0.0),
            (1 / 4, 0.0, 3 / 4, 0.0, 0.0), (1 / 4, 0.0, 3 / 4, 0.0, 0.0))

    def _dirk_weights(self, dt: float):
        """The right-hand-side weights of the four stages for step size dt, laid
        out for the stage plans: per stage [1, bimp_0, bexp_0, bexp_1, wimp_lin0
        BuoFac, -, -, -, bimp_1, bexp_2, bimp_2, bimp_3] (BPR353.set_weights,
        time_scheme.py:230-238; wimp_lin0 = dt / 2)."""
        w = self.w_stage_host
        w.zero_()
        for k in range(1, 5):
            B = [v * dt for v in self.BIMP[k]]
            E = [v * dt for v in self.BEXP[k]]
            o = 16 * (k - 1)
            w[o + 0], w[o + 1], w[o + 2] = 1.0, B[0], E[0]
            w[o + 3], w[o + 8] = E[1], B[1]
            w[o + 9], w[o + 10], w[o + 11] = E[2], B[2], B[3]
            w[o + 4] = 0.5 * dt * self.p.BuoFac
        self.wdev_stage.copy_(w, non_blocking=True)

    def _one_step_dirk(self, n_time_step: int, dt: float) -> float:
        """BPR353: stages 1-3 run the radial loop into explicit slot k-1; stage 1
        also takes the dt decision (this scheme's Courant factors) and rebuilds
        the operators for wimp_lin0 = dt/2 on the first step or a new dt; every
        stage runs the lm loop (step_time.py:1640-1680)."""
        self._check_info()
        dt_new = dt
        for k in range(1, 5):
            self.stage = k
            if k <= 3:
                self.expl_cur = self.explS[k - 1]
                self._exec("radial")
                if k == 1:
                    self._cfl_copy()
                    dtrkc, dthkc = self._cfl_values()
                    l_new_dt, dt_new = dt_courant(dt, self.dtMax, dtrkc, dthkc)
                    if self.dt_schedule is not None and n_time_step <= len(self.dt_schedule):
                        dt_new = float(self.dt_schedule[n_time_step - 1])
                        l_new_dt = dt_new != dt
                    self.dt_log.append(dt_new)
                    self.n_step = n_time_step
                    self.dt_hist = [dt_new, dt_new]
                    self._dirk_weights(dt_new)
                    if n_time_step == 1 or l_new_dt or self.wimp_lin0 is None:
                        self.build_all_matrices(0.5 * dt_new)
            self._exec("lm")
        self.expl_cur = self.explS[0]
        self.time += dt_new
        return dt_new

    def begin_step(self):
        """Queue the radial loop and the copy of the Courant maxima (no host wait)."""
        self._exec("radial")
        self.expl_pending = True
        self._cfl_copy()

    def finish_step(self, n_time_step: int, dt: float) -> float:
        """Wait for the Courant maxima, decide dt, rebuild if needed, queue the lm loop."""
        self._check_info()  # a deferred check of the previous rebuild (done by now)
        dtrkc, dthkc = self._cfl_values()
        l_new_dt, dt_new = dt_courant(dt, self.dtMax, dtrkc, dthkc)
        if self.dt_schedule is not None and n_time_step <= len(self.dt_schedule):
            # replay another run's step sizes (precision comparisons on one dt sequence)
            dt_new = float(self.dt_schedule[n_time_step - 1])
            l_new_dt = dt_new != dt
        self.dt_log.append(dt_new)
        self.n_step = n_time_step
        self.dt_hist = [dt_new, self.dt_hist[0]]
        # the reference's rule (step_time.py:1628-1629): rebuild on the first
        # step or when the Courant decision changed dt; a run whose operators
        # were never built (the reference would fail there) builds them once
        if n_time_step == 1 or l_new_dt or self.wimp_lin0 is None:
            self.build_all_matrices(self.p.alpha * dt_new)
        self.lm_loop()
        self.time += dt_new
        return dt_new

    # ------------------------------------------------------------ device step control
    LOG_CAP = 1024  # dt log ring on the device; the host reads it at least this often

    def _ctl_buffers(self):
        """Device state of the step control, the Courant spacings, the dt log
        ring and the replayed step sizes (created once; re-uploaded in place)."""
        dev, rad = self.device, self.res.rad
        c = self._ctl
        if c is None:
            up = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=F64).to(dev)
            c = self._ctl = dict(state=torch.zeros(C.sizeof(_lib.StepState), dtype=torch.uint8,
                                                   device=dev),
                                 delxr2=up(rad["delxr2"]), delxh2=up(rad["delxh2"]),
                                 log=torch.zeros(self.LOG_CAP, dtype=F64, device=dev),
                                 sched=None, n_sched=0, pulled=0)
        sch = self.dt_schedule
        n = 0 if sch is None else len(sch)
        if n != c["n_sched"]:
            # the schedule's length is a launch argument: new plans and graphs
            c["sched"] = torch.zeros(max(n, 1), dtype=F64, device=dev)
            c["n_sched"] = n
            for key in [k for k in self._plans if k[0] == "step"]:
                self._plans.pop(key)
                self._graphs.pop(key, None)
        if n:
            c["sched"].copy_(torch.as_tensor(np.asarray(sch, dtype=np.float64)))
        return c

    def _ctl_launch(self):
        c = self._ctl
        a = _lib.StepCtlArgs()
        a.cfl, a.delxr2, a.delxh2 = ptr(self.cfl), ptr(c["delxr2"]), ptr(c["delxh2"])
        a.state, a.w = ptr(c["state"]), ptr(self.wdev)
        if c["n_sched"]:
            a.dt_schedule, a.n_schedule = ptr(c["sched"]), c["n_sched"]
        a.dt_log, a.log_cap = ptr(c["log"]), self.LOG_CAP
        a.info, a.info_rows, a.info_cols = ptr(self.gj_info), self.gj_info.shape[0], self.gj_info.shape[1]
        a.N, a.alpha, a.BuoFac = self.N, self.p.alpha, self.p.BuoFac
        if self.dirk:
            a.w_stage, a.scheme = ptr(self.wdev_stage), 1
        with self._t("step_ctl"):
            _lib.call("mh_step_control", _lib.ref(a), stream_handle())

    def _step_launches(self):
        """A whole step as one launch sequence: radial loop, device step control,
        the rebuild launches (no-ops unless the step control asks for a
        rebuild) and the lm loop."""
        base = ptr(self._ctl["state"])
        ctl = (base + _lib.StepState.wimp_lin0.offset, base + _lib.StepState.rebuild.offset)
        if self.dirk:
            # BPR353: radial loops of stages 1-3 (explicit slots 0-2), the decision
            # and the rebuild after the first, the lm loop of every stage
            for k in range(1, 5):
                self.stage = k
                if k <= 3:
                    self.expl_cur = self.explS[k - 1]
                    self._radial_launches()
                    if k == 1:
                        self._ctl_launch()
                        with self._t("rebuild"):
                            self._rebuild_launches(0.0, ctl)
                self._lm_launches()
            self.expl_cur = self.explS[0]
            return
        self._radial_launches()
        self._ctl_launch()
        with self._t("rebuild"):
            self._rebuild_launches(0.0, ctl)
        self._lm_launches()

    def _ctl_push(self, dt):
        """Upload the host's stepping state (before device-controlled steps)."""
        c = self._ctl_buffers()
        st = _lib.StepState()
        st.dt = float(dt)
        st.dtMax = self.dtMax
        st.dt1, st.dt2 = self.dt_hist
        st.wimp_lin0 = math.nan if self.wimp_lin0 is None else self.wimp_lin0
        st.time = self.time
        st.n_step, st.n_rebuild = self.n_step, self.n_rebuild
        st.rebuild = 0
        host = torch.frombuffer(bytearray(bytes(st)), dtype=torch.uint8)
        c["state"].copy_(host)   # synchronous upload from pageable memory
        c["pulled"] = self.n_step

    def _ctl_pull(self):
        """Read the device stepping state and the dt log back (synchronises the
        stream), check the pivots, and return the dt of the last step."""
        c = self._ctl
        st = _lib.StepState.from_buffer_copy(c["state"].cpu().numpy().tobytes())
        n_new = st.n_step - c["pulled"]
        if n_new > self.LOG_CAP:
            raise RuntimeError("dt log overrun: read the device state at least every "
                               f"{self.LOG_CAP} steps")
        log = c["log"].cpu().numpy()
        for n in range(c["pulled"] + 1, st.n_step + 1):
            self.dt_log.append(float(log[(n - 1) % self.LOG_CAP]))
        c["pulled"] = st.n_step
        self.n_step, self.n_rebuild, self.time = st.n_step, st.n_rebuild, st.time
        self.dt_hist = [st.dt1, st.dt2]
        self.wimp_lin0 = st.wimp_lin0
        if st.sing_step:
            name = _INFO_ROWS[st.sing_sys]
            raise AssertionError(f"Singular {name} matrix for l={st.sing_l}, info={st.sing_info} "
                                 f"(rebuild of step {st.sing_step})")
        if st.rebuild:  # the last step's rebuild has not been scanned on the device
            self._info_host.copy_(self.gj_info, non_blocking=True)
            if self._info_event is None:
                self._info_event = torch.cuda.Event()
            self._info_event.record()
            self._info_pending = True
            self._check_info()
        return st.dt

    def step_async(self):
        """Queue one whole step with the decisions on the device (one HIP graph
        when use_graphs); no host wait.  Call ctl_begin(dt) once before and
        ctl_end() to bring the stepping state back to the host."""
        self._exec("step")
        if not self.dirk:   # CNAB2: the explicit-history buffers swap roles
            self.expl_cur, self.expl_prev = self.expl_prev, self.expl_cur
            self._par ^= 1
        self._ctl["queued"] += 1
        if self._ctl["queued"] - self._ctl["pulled"] >= self.LOG_CAP:
            self._ctl_pull()

    def save_checkpoint(self, path, step: int | None = None, sim_time: float | None = None):
        """Restart file in the reference's .pt layout (``main.py:189-223``)."""
        from .checkpoint import save_checkpoint
        save_checkpoint(self, path, step, sim_time)

    def load_checkpoint(self, path):
        """Restore from a restart file (``main.py:226-267``); returns (step, sim_time).
        Continue with one_step(step + 1, self.dt_hist[0]) or run_steps."""
        from .checkpoint import load_checkpoint
        return load_checkpoint(self, path)

    def ctl_begin(self, dt: float):
        """Start device-controlled stepping; dt: the dt passed to the first
        step's decision, as for one_step."""
        self._check_info()
        self._ctl_push(dt)
        self._ctl["queued"] = self.n_step

    def ctl_end(self) -> float:
        """Finish device-controlled stepping: read the state back (dt_log, time,
        n_rebuild, dt history), check the pivots; returns the last step's dt."""
        return self._ctl_pull()

    def run_steps(self, n_steps: int, dt: float) -> float:
        """Advance n_steps steps with every decision on the device: each step is
        one launch sequence (one HIP graph replay when use_graphs) and the host
        never waits inside the loop.  Same decisions, weights and results as
        n_steps calls of one_step; dt as for one_step's first call.  Returns the
        dt of the last step (CNAB2 or BPR353)."""
        self.ctl_begin(dt)
        for _ in range(n_steps):
            self.step_async()
        return self.ctl_end()

    # ------------------------------------------------------------ diagnostics
    def field_set(self, **given):
        """{w, dw, z, b, db, aj}: the run's own field tensors, with any given
        (lm_max, n_r_max) arrays in their place (moved to the run's device as
        complex128 when they are not there already).  The diagnostics below
        compute on these, as the reference's do on the tensors passed to them
        (``output.py:111-193,750-861``)."""
        fs = {k: self.f[k] for k in ("w", "dw", "z", "b", "db", "aj")}
        shape = tuple(self.f["w"].shape)
        for k, v in given.items():
            if v is None:
                continue
            if k not in fs:
                raise KeyError(k)
            t = torch.as_tensor(v)
            if tuple(t.shape) != shape:
                raise ValueError(f"{k}: expected shape {shape}, got {tuple(t.shape)}")
            fs[k] = t.to(device=self.device, dtype=C128).contiguous()
        return fs

    def _energy_out(self, by_m: int, fields=None, LFfac=None):
        lp1 = self.p.l_max + 1
        shape = {0: (4, lp1), 1: (4, lp1), 2: (16, lp1), 3: (8, lp1, self.N)}[by_m]
        out = torch.zeros(shape, dtype=F64, device=self.device)
        a = _lib.EnergyArgs()
        f = fields if fields is not None else self.f
        a.w, a.dw, a.z, a.b, a.db, a.aj = (ptr(f[k]) for k in ("w", "dw", "z", "b", "db", "aj"))
        a.lm_start = ptr(self.res.d["lm_start"])
        a.or2, a.qw, a.out = ptr(self.res.d["or2"]), ptr(self.res.d["q"]), ptr(out)
        a.l_max, a.N, a.by_m, a.l_mag = self.p.l_max, self.N, int(by_m), int(self.p.l_mag)
        a.minc = self.p.minc
        a.LFfac = self.p.LFfac if LFfac is None else LFfac
        if self.anel:
            a.orho1 = ptr(self.d_an["orho1"])
        _lib.call("mh_energy", _lib.ref(a), stream_handle())
        return out

    def energies(self, fields=None):
        """(e_kin_pol, e_kin_tor, e_mag_pol, e_mag_tor) (``output.py:111-193``)."""
        e = self._energy_out(0, fields).cpu().numpy()
        return tuple(float(x) for x in e.sum(axis=1))

    def energy_components(self, fields=None):
        """All components of the reference's e_kin.TAG (8) and e_mag_oc.TAG (12)
        lines (``output.py:111-193``): totals, axisymmetric (m = 0), equatorially
        symmetric ((l+m) even for kinetic poloidal and magnetic toroidal, odd for
        the other two), their m = 0 parts, and the poloidal magnetic energy of the
        potential field outside the shell from the CMB values.  Returns
        (ekin, emag) as tuples in the reference's column order."""
        f = fields if fields is not None else self.f
        e = self._energy_out(2, f).cpu().numpy().reshape(4, 4, -1).sum(axis=2)  # [q][bin]
        # bins: 0 m>0 (l+m) even, 1 m>0 odd, 2 m=0 l even, 3 m=0 l odd
        tot = e.sum(axis=1)
        as_ = e[:, 2] + e[:, 3]
        even, odd = e[:, 0] + e[:, 2], e[:, 1] + e[:, 3]
        ekin = (tot[0], tot[1], as_[0], as_[1], even[0], odd[1], e[0, 2], e[1, 3])
        if self.p.l_mag:
            r, maps = self.res, self.res.maps
            l = np.asarray(maps["l"], dtype=np.float64)
            m0 = np.asarray(maps["m"]) == 0
            bc = f["b"][:, 0].cpu().numpy()  # radial index 0 is the CMB
            cc = np.where(m0, bc.real ** 2, 2.0 * (bc.real ** 2 + bc.imag ** 2))
            w = 0.5 * self.p.LFfac / r.rad["r_cmb"] * l * l * (l + 1.0) * cc
            e_os, e_as_os = float(w.sum()), float(w[m0].sum())
            emag = (tot[2], tot[3], as_[2], as_[3], e_os, e_as_os, odd[2], even[3], e[2, 3],
                    e[3, 2], 0.0, 0.0)
        else:
            emag = (0.0,) * 12
        return tuple(float(x) for x in ekin), tuple(float(x) for x in emag)

    def spectra(self, fields=None):
        """Per-degree and per-order kinetic and magnetic energy spectra: the
        radially integrated energies of each degree l and of each order m (the
        per-l and per-m aggregation of ``get_dlm``, ``output.py:768-780``, with
        the 1/2 and LFfac of ``get_e_kin_full`` / ``get_e_mag_oc_full``); each
        sums to the total energy."""
        el = self._energy_out(0, fields).cpu().numpy()
        em = self._energy_out(1, fields).cpu().numpy()
        return dict(kin_l=el[0] + el[1], mag_l=el[2] + el[3], kin_m=em[0] + em[1],
                    mag_m=em[2] + em[3])

    def get_dlm(self, switch="V", fields=None):
        """Integral length scales of the flow (switch 'V': w, dw, z) or of the
        field ('B': b, db, aj), as ``get_dlm`` (``output.py:750-861``):
        (dl, dlR, dm, dlc, dlPolPeak, dlRc, dlPolPeakR); dlR, dlRc and
        dlPolPeakR are numpy arrays over the radial levels.  The per-(l, r)
        profiles and the per-order integrals come from the device reduction
        (``mh_energy`` modes 3 and 1, the field's LFfac left out as the
        reference does); the few (l_max+1) x n_r_max sums are formed here."""
        if switch not in ("V", "B"):
            raise ValueError(f"switch must be 'V' or 'B', not {switch!r}")
        f = fields if fields is not None else self.f
        if switch == "B":
            f = dict(f, w=f["b"], dw=f["db"], z=f["aj"])   # the kinetic slots carry b
        prof = self._energy_out(3, f, LFfac=1.0)[:4].cpu().numpy()   # halved, per (l, r)
        per_m = self._energy_out(1, f, LFfac=1.0)[:2].cpu().numpy().sum(axis=0)
        q = self.res.rad["q"]
        l_max, N = self.p.l_max, self.N
        l_first = 1 if switch == "V" else 2
        e_lr = prof[0] + prof[1]           # 0.5 e_lr of output.py:768-772
        e_pol_lr = prof[0]
        e_lr_c = prof[2] + prof[3]
        ls = np.arange(l_first, l_max + 1, dtype=np.float64)
        E_l = e_lr[l_first:] @ q
        E, EL = float(E_l.sum()), float((ls * E_l).sum())
        dl = math.pi * E / EL if EL != 0.0 else 0.0
        E_pol = np.zeros(l_max + 1)
        E_pol[l_first:] = e_pol_lr[l_first:] @ q
        lpeak = int(np.argmax(E_pol[1:])) + 1
        dlPolPeak = math.pi / lpeak
        eps = np.finfo(np.float64).eps
        if switch == "V":
            Ec_l = e_lr_c[l_first:] @ q
            Ec, ELc = float(Ec_l.sum()), float((ls * Ec_l).sum())
            dlc = math.pi * Ec / ELc if ELc != 0.0 else 0.0
        else:
            dlc = 0.0
        ms = np.arange(self.p.minc, l_max + 1, self.p.minc)
        E_m, EM_m = float(per_m[ms].sum()), float((ms * per_m[ms]).sum())
        dm = math.pi * E_m / EM_m if EM_m != 0.0 else 0.0
        if switch == "V":
            def scale(e):
                er = e[l_first:].sum(axis=0)
                elr = (ls[:, None] * e[l_first:]).sum(axis=0)
                return np.where(elr > 10.0 * eps, math.pi * er / np.where(elr != 0, elr, 1.0), 0.0)
            dlR, dlRc = scale(e_lr), scale(e_lr_c)
            per_r = 2.0 * e_pol_lr[1:]      # the reference compares the unhalved values
            mx, ix = per_r.max(axis=0), per_r.argmax(axis=0) + 1
            dlPolPeakR = np.where(mx > 10.0 * eps, math.pi / ix, 0.0)
        else:
            dlR, dlRc, dlPolPeakR = np.zeros(N), np.zeros(N), np.zeros(N)
        return dl, dlR, dm, dlc, dlPolPeak, dlRc, dlPolPeakR


def dt_courant(dt, dtMax, dtrkc_min, dthkc_min):
    """Hysteresis dt decision (reference ``courant.py:130-155``)."""
    dt_fac = 2.0
    dtMin = min(dtrkc_min, dthkc_min)
    dt_2 = min(0.5 * (1.0 / dt_fac + 1.0) * dtMin, dtMax)
    if dt > dtMax:
        return True, dtMax
    if dt > dtMin:
        return True, dt_2
    if dt_fac * dt < dtMin and dt < dtMax:
        return True, dt_2
    return False, dt
