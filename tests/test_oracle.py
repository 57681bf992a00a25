"""Pins of the CPU oracle: golden vectors, closed forms, library routines, invariants.

A mistake in the oracle (a dropped term, a wrong sign or index, a transposed
operand) must fail one of these.  Nothing here retypes an oracle formula:
every expected value comes from the reference's own trajectory fixtures, an
independent library routine (numpy / scipy / LAPACK), a closed form, or brute
force.
"""

import json
import math
import os

import numpy as np
import pytest
import scipy.fft
import scipy.linalg
import scipy.special

from oracle import algebra, horizontal, matrices, radial
from oracle.dynamo import OracleRun
from oracle.params import Params
from oracle.sht import SHT

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ------------------------------------------------------------ golden trajectory
# l=0 pressure of an anelastic run: its system carries the integral condition
# and has a condition number near 3e9 (strat = 5), and p(l=0) ~ 1e8 from the
# hydrostatic balance; numpy's inverse and the reference's (torch/MKL) differ by
# 3e-10 there (torch's own inverse reproduces the fixture to 2e-14)
P_TOL_ANEL = 1e-9


@pytest.mark.parametrize("name", ["bench16", "c1", "dtchange", "bpr16", "bprdt", "minc2",
                                  "bprminc4", "condic", "bprcondic", "rotic", "bprrotic4",
                                  "anel21", "anelmag16", "bpranel2", "amc16", "rotom8",
                                  "fd2mhd", "fd4mhd", "fdanel", "fdanelmhd", "fd4anel",
                                  "fdbpr", "fdcondic", "fdrotic"])
def test_oracle_matches_reference_trajectory(name):
    """Energies every step and fields at the saved steps, from the reference's
    one_step with the returned dt passed to the next step; the dt sequence
    itself (dtchange: five Courant-driven changes in 30 steps, each with a
    rebuild and variable-step CNAB2 weights; bprdt: the same with the BPR353
    four-stage DIRK, whose decision is taken at stage 1 with its own Courant
    factors) is matched step by step.  bpr16: BPR353 at constant dt.  minc2 and
    bprminc4: azimuthal symmetry (orders m = 0, minc, 2 minc, ... on one
    2 pi / minc sector) with CNAB2 and with BPR353.  condic and bprcondic: a
    conducting inner core (coupled outer/inner B and J solves), whose fields
    b_ic and aj_ic are matched too.  rotic and bprrotic4: the core rotating,
    with its rotation rate omega_ic matched every step.  anel21, anelmag16 and
    bpranel2: the anelastic (polytropic) step: the hydrodynamic benchmark's
    physics (strat 5, stress-free walls, angular-momentum corrections), a
    magnetic run (strat 1, gravity ~ r) and BPR353 with minc = 2; the pressure
    within P_TOL_ANEL.  amc16: Boussinesq, stress-free, with both
    angular-momentum corrections.  rotom8: a rotating core started at
    omega_ic1 = 2.  fd2mhd, fd4mhd, fdanel, fdanelmhd, fdbpr: the
    finite-difference radial scheme with the double-curl poloidal equation
    (second- and fourth-order stencils on the stretched grid, an even n_r with
    Simpson's even-N rule in the energies, anelastic hydrodynamic and magnetic
    runs, BPR353)."""
    case = json.load(open(os.path.join(GOLD, f"traj_{name}.json")))
    g = np.load(os.path.join(GOLD, f"traj_{name}.npz"))
    run = OracleRun(case["cfg"])
    run.initialize_fields()
    if case["noise"]:
        run.add_noise(case["noise"]["seed"], case["noise"]["amp"])
    run.setup_initial_state()
    run.initialize_dt(case["dt"])
    dt = case["dt"]
    for n in range(1, case["steps"] + 1):
        dt = run.one_step(n, dt)
        assert dt == pytest.approx(g["dt"][n - 1], rel=1e-12), (n, dt, g["dt"][n - 1])
        e = np.array(run.energies())
        assert np.all(np.abs(e - g["energies"][n - 1]) <= 1e-11 * np.abs(g["energies"][n - 1]))
        if "omega_ic" in g.files:
            assert run.omega_ic == pytest.approx(g["omega_ic"][n - 1], rel=1e-11), n
        if n in case["field_steps"]:
            keys = ("s", "z", "w", "p", "b", "aj") + (("b_ic", "aj_ic") if run.cond_ic else ())
            for k in keys:
                ref = g[f"{k}_step{n}"]
                tol = P_TOL_ANEL if (k == "p" and run.anel) else 1e-11
                assert np.abs(getattr(run, k) - ref).max() <= tol * np.abs(ref).max(), (n, k)
    if name in ("dtchange", "bprdt"):   # the fixture does take the Courant branch
        assert len(set(g["dt"].tolist())) >= {"dtchange": 5, "bprdt": 4}[name]


def test_oracle_step1_kat_from_reference_log():
    """Dev-log known answer (PROGRESS.md:90): step-1 e_kin_pol=4.446e-01, e_kin_tor=2.218e+00."""
    run = OracleRun({"l_max": 16, "n_r_max": 33})
    run.initialize_fields()
    run.setup_initial_state()
    run.initialize_dt(1e-4)
    run.one_step(1, 1e-4)
    ekp, ekt, _, _ = run.energies()
    assert f"{ekp:.3e}" == "4.446e-01" and f"{ekt:.3e}" == "2.218e+00"


# ------------------------------------------------------------ horizontal grid
@pytest.mark.parametrize("n", [8, 48, 96])
def test_gauleg_against_numpy(n):
    th, w = horizontal.gauleg(n)
    x, wn = np.polynomial.legendre.leggauss(n)
    assert np.abs(np.cos(th) - x[::-1]).max() < 1e-14
    assert np.abs(w - wn[::-1]).max() < 1e-14


def test_gauleg_exact_orthogonality_n576():
    """n=576 nodes integrate P_l P_l' exactly (l + l' < 2n): numpy's leggauss is only
    documented to degree 100, so large n is pinned by the quadrature property itself."""
    th, w = horizontal.gauleg(576)
    x = np.cos(th)
    ls = np.arange(0, 576, 23)
    P = np.array([scipy.special.eval_legendre(l, x) for l in ls])
    G = (P * w[None, :]) @ P.T
    assert np.abs(G - np.diag(2.0 / (2 * ls + 1))).max() < 1e-13


def _plm_scipy(l, m, theta):
    """Orthonormal P_lm without Condon-Shortley phase from scipy's lpmv."""
    norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.exp(math.lgamma(l - m + 1) - math.lgamma(l + m + 1)))
    return norm * (-1) ** m * scipy.special.lpmv(m, l, np.cos(theta))


def test_plm_against_scipy_and_closed_forms():
    theta = np.array([0.1, 0.7, 1.3, 2.0, 2.9])
    l_max = 20
    P, dP = horizontal.plm_theta(theta, l_max)
    lm = 0
    for m in range(l_max + 1):
        for l in range(m, l_max + 1):
            ref = _plm_scipy(l, m, theta)
            assert np.abs(P[lm] - ref).max() < 1e-12 * max(1.0, np.abs(ref).max()), (l, m)
            # sin(theta) dP/dtheta by a centred difference of the scipy values
            h = 1e-5
            fd = np.sin(theta) * (_plm_scipy(l, m, theta + h) - _plm_scipy(l, m, theta - h)) / (2 * h)
            assert np.abs(dP[lm] - fd).max() < 1e-6 * max(1.0, np.abs(fd).max()), (l, m)
            lm += 1
    assert np.allclose(P[0], 1.0 / math.sqrt(4 * math.pi))
    assert np.allclose(P[1], math.sqrt(3.0 / (4 * math.pi)) * np.cos(theta))


def test_plm_orthonormal_with_gauss_weights():
    p = Params({"l_max": 24})
    grid = horizontal.grid_theta(p)
    th, w = grid["theta_ord"], grid["gauss"]
    P, _ = horizontal.plm_theta(th, p.l_max)
    G = (P * w[None, :]) @ P.T * 2 * math.pi
    maps = horizontal.lm_maps(p)
    same_m = maps["lm2m"][:, None] == maps["lm2m"][None, :]
    assert np.abs(G[same_m] - np.eye(p.lm_max)[same_m]).max() < 1e-13


# ------------------------------------------------------------ radial
def test_chebyshev_against_cosines_and_polynomial_derivatives():
    p = Params({"l_max": 8, "n_r_max": 25})
    r, x, r_cmb, r_icb = radial.cheb_grid(p)
    R, D, D2, D3 = radial.cheb_mats(p, x, r_cmb, r_icb)
    n = np.arange(25)
    assert np.abs(R - np.cos(np.outer(np.arccos(np.clip(x, -1, 1)), n))).max() < 1e-12
    D1m, D2m, D3m = radial.diff_mats(p, (R, D, D2, D3))
    f = r ** 5 - 2 * r ** 3 + 1.0 / 7.0
    assert np.abs(D1m @ f - (5 * r ** 4 - 6 * r ** 2)).max() < 1e-10
    assert np.abs(D2m @ f - (20 * r ** 3 - 12 * r)).max() < 1e-9
    assert np.abs(D3m @ f - (60 * r ** 2 - 12)).max() < 1e-7


def test_costf_against_scipy_dct_and_involution():
    rng = np.random.default_rng(0)
    f = rng.standard_normal((3, 33))
    ref = scipy.fft.dct(f, type=1, axis=-1) / math.sqrt(2 * 32)
    assert np.abs(radial.costf(f) - ref).max() < 1e-13
    assert np.abs(radial.costf(radial.costf(f)) - f).max() < 1e-13


def test_rint_exact_on_polynomials():
    p = Params({"l_max": 8, "n_r_max": 17})
    rad = radial.radial_functions(p)
    q = radial.rint_weights(p, rad["r_cmb"], rad["r_icb"])
    r, a, b = rad["r"], rad["r_icb"], rad["r_cmb"]
    for k in range(0, 10):
        assert abs(radial.rint(r ** k, q) - (b ** (k + 1) - a ** (k + 1)) / (k + 1)) < 1e-12 * b ** (k + 1)


@pytest.mark.parametrize("n", [9, 10, 33, 40])
def test_simpson_weights_exact_on_low_degree(n):
    """simps is exact for linear functions on any grid (both parities) and for
    quadratics with an odd number of points (three-point panels); checked on
    the stretched finite-difference grid against the closed-form integrals."""
    from oracle import fd
    r = fd.fd_grid(n, 0.5, 1.5, 0.3, 0.1)
    q = radial.simps_weights(r)
    a, b = r[-1], r[0]
    assert q @ (2.0 + 3.0 * r) == pytest.approx(2.0 * (b - a) + 1.5 * (b * b - a * a), rel=1e-14)
    quad = q @ (r * r)
    exact = (b ** 3 - a ** 3) / 3.0
    if n % 2:
        assert quad == pytest.approx(exact, rel=1e-14)
    else:
        assert quad == pytest.approx(exact, rel=1e-3) and quad != pytest.approx(exact, rel=1e-13)


def test_conduction_state_analytic():
    """Boussinesq s0 = A + B/r with s0(r_cmb)=0, s0(r_icb)=sqrt(4 pi); p0' = BuoFac (r/r_cmb) s0."""
    p = Params({"l_max": 8, "n_r_max": 33, "ra": 1e5})
    rad = radial.radial_functions(p)
    s0, p0 = matrices.ps_cond(p, rad, radial.cheb_mats(p))
    r, ri, ro = rad["r"], rad["r_icb"], rad["r_cmb"]
    B = math.sqrt(4 * math.pi) / (1 / ri - 1 / ro)
    A = -B / ro
    assert np.abs(s0 - (A + B / r)).max() < 1e-12
    P = lambda x: p.BuoFac / ro * (A * x * x / 2 + B * x)
    assert np.abs(p0 - (P(r) - P(ro))).max() < 1e-9 * np.abs(p0).max()


# ------------------------------------------------------------ SHT
def _brute_synthesis(p, S, theta, phi):
    """sum_lm P_lm(theta) S_lm e^{i m phi} with the real-field convention (scipy P_lm)."""
    maps = horizontal.lm_maps(p)
    out = np.zeros((theta.size, phi.size))
    for lm in range(p.lm_max):
        l, m = maps["lm2l"][lm], maps["lm2m"][lm]
        P = _plm_scipy(l, m, theta)[:, None]
        if m == 0:
            out += P * S[lm].real
        else:
            out += 2.0 * P * (S[lm] * np.exp(1j * m * phi))[None, :].real
    return out


def _rand(p, B, seed):
    maps = horizontal.lm_maps(p)
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((p.lm_max, B)) + 1j * rng.standard_normal((p.lm_max, B))
    x[maps["lm2m"] == 0] = x[maps["lm2m"] == 0].real
    return x


def test_scal_to_spat_brute_force():
    p = Params({"l_max": 8})
    sht = SHT(p)
    S = _rand(p, 1, 1)[:, 0]
    g = sht.scal_to_spat(S)
    theta = sht.grid["theta_grid"]
    phi = 2 * math.pi * np.arange(p.n_phi_max) / p.n_phi_max
    assert np.abs(g - _brute_synthesis(p, S, theta, phi)).max() < 1e-12


def test_scal_to_spat_brute_force_minc():
    """minc = 3: the sector's longitudes phi_k = 2 pi k / (minc n_phi_max) and
    the orders m = 0, 3, 6, ... of the lm order; sector periodicity follows."""
    p = Params({"l_max": 12, "minc": 3})
    assert p.n_phi_tot == 36 and p.n_phi_max == 12 and p.n_theta_max == 18
    assert p.lm_max == sum(13 - m for m in range(0, 13, 3))
    sht = SHT(p)
    S = _rand(p, 1, 5)[:, 0]
    g = sht.scal_to_spat(S)
    theta = sht.grid["theta_grid"]
    phi = 2 * math.pi * np.arange(p.n_phi_max) / p.n_phi_tot
    assert np.abs(g - _brute_synthesis(p, S, theta, phi)).max() < 1e-12
    assert np.abs(sht.scal_to_SH(g) - S).max() < 1e-12


def test_sht_round_trips():
    p = Params({"l_max": 20})
    sht = SHT(p)
    S, T = _rand(p, 3, 2), _rand(p, 3, 3)
    assert np.abs(sht.scal_to_SH(sht.scal_to_spat(S)) - S).max() < 1e-12
    S[0] = T[0] = 0.0
    _, vt, vp = sht.torpol_to_spat(np.zeros_like(S), S, T)
    S2, T2 = sht.spat_to_sphertor(vt, vp)
    assert np.abs(S2 - S).max() < 1e-12 and np.abs(T2 - T).max() < 1e-12


def test_torpol_to_spat_is_sin_weighted_gradient_and_curl():
    """S only: (sin th d/dth f, d/dphi f); T only: (1/sin th... ) Robert-form curl of f = scal_to_spat."""
    p = Params({"l_max": 12})   # n_theta must be even: l_max=10 gives 15 latitudes
    sht = SHT(p)
    maps = horizontal.lm_maps(p)
    S = _rand(p, 1, 4)[:, 0]
    S[0] = 0.0
    zero = np.zeros_like(S)
    br, bt, bp = sht.torpol_to_spat(zero, S, zero)
    theta = sht.grid["theta_grid"]
    phi = 2 * math.pi * np.arange(p.n_phi_max) / p.n_phi_max
    h = 1e-5
    dfdth = (_brute_synthesis(p, S, theta + h, phi) - _brute_synthesis(p, S, theta - h, phi)) / (2 * h)
    assert np.abs(br).max() < 1e-13
    assert np.abs(bt - np.sin(theta)[:, None] * dfdth).max() < 1e-6 * np.abs(bt).max()
    dfdphi = sht.scal_to_spat(1j * maps["lm2m"] * S)
    assert np.abs(bp - dfdphi).max() < 1e-12 * np.abs(bp).max()
    # toroidal: (d/dphi f, -sin th d/dth f)
    _, tt, tp = sht.torpol_to_spat(zero, zero, S)
    assert np.abs(tt - dfdphi).max() < 1e-12 * np.abs(tt).max()
    assert np.abs(tp + np.sin(theta)[:, None] * dfdth).max() < 1e-6 * np.abs(tp).max()
    # radial part: Q synthesises like a scalar
    qr, _, _ = sht.torpol_to_spat(S, zero, zero)
    assert np.abs(qr - sht.scal_to_spat(S)).max() < 1e-13


# ------------------------------------------------------------ algebra
def _rand_band(n, kl, ku, seed=42):
    rng = np.random.default_rng(seed)
    A = np.zeros((n, n))
    for j in range(n):
        for i in range(max(0, j - ku), min(n, j + kl + 1)):
            A[i, j] = rng.standard_normal()
    A[np.arange(n), np.arange(n)] = np.abs(A).sum(axis=1) + 1.0
    return A


@pytest.mark.parametrize("n,kl,ku", [(5, 1, 1), (10, 1, 1), (10, 2, 2), (33, 2, 2), (73, 2, 2), (64, 3, 3)])
def test_band_lu_round_trip_real_and_complex(n, kl, ku):
    """Restates the reference's TestPivotedBandLU round trips (tests/test_banded_solvers.py:290-325)."""
    A = _rand_band(n, kl, ku)
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    abd, piv, info = algebra.prepare_band(algebra.dense_to_band_storage(A, kl, ku), n, kl, ku)
    assert info == 0
    assert np.abs(algebra.solve_band(abd, n, kl, ku, piv, A @ x) - x).max() < 1e-11
    assert np.abs(algebra.solve_band(abd, n, kl, ku, piv, A @ x.real) - x.real).max() < 1e-11


def _general_band(n, kl, ku, seed):
    rng = np.random.default_rng(seed)
    A = np.zeros((n, n))
    for j in range(n):
        for i in range(max(0, j - ku), min(n, j + kl + 1)):
            A[i, j] = rng.standard_normal()
    return A + 0.1 * np.eye(n)


@pytest.mark.parametrize("n,kl,ku,seed", [(12, 1, 1, 0), (40, 2, 2, 1), (60, 3, 3, 2), (30, 2, 3, 3)])
def test_band_pivots_and_solution_match_lapack_dgbtrf(n, kl, ku, seed):
    """Same first-maximum partial pivoting as LAPACK dgbtrf: identical pivots, same solution."""
    A = _general_band(n, kl, ku, seed)
    abd, piv, info = algebra.prepare_band(algebra.dense_to_band_storage(A, kl, ku), n, kl, ku)
    assert info == 0
    ab = np.zeros((2 * kl + ku + 1, n))
    for j in range(n):
        for i in range(max(0, j - ku), min(n, j + kl + 1)):
            ab[kl + ku + i - j, j] = A[i, j]
    lub, ipiv, linfo = scipy.linalg.lapack.dgbtrf(ab, kl, ku)
    assert linfo == 0
    assert np.array_equal(piv[:-1], ipiv[:-1] + 1)
    b = np.random.default_rng(seed + 10).standard_normal(n)
    x = algebra.solve_band(abd, n, kl, ku, piv, b)
    assert np.abs(x - np.linalg.solve(A, b)).max() < 1e-10 * max(1.0, np.abs(x).max())


def test_band_forced_pivoting_and_singular_info():
    """Reference case with pivoting required (tests/test_banded_solvers.py:345-372) + singular info."""
    n = 10
    A = np.zeros((n, n))
    A[np.arange(n), np.arange(n)] = 0.01
    A[np.arange(1, n), np.arange(n - 1)] = 5.0
    A[np.arange(n - 1), np.arange(1, n)] = 0.1
    A[0, 0] = A[-1, -1] = 10.0
    A = A + 2.0 * np.eye(n)
    b = np.random.default_rng(0).standard_normal(n)
    abd, piv, info = algebra.prepare_band(algebra.dense_to_band_storage(A, 1, 1), n, 1, 1)
    assert info == 0 and np.any(piv[:-1] != np.arange(1, n))
    assert np.abs(A @ algebra.solve_band(abd, n, 1, 1, piv, b) - b).max() < 1e-10
    S = A.copy()
    S[:, 3] = 0.0
    _, _, info = algebra.prepare_band(algebra.dense_to_band_storage(S, 1, 1), n, 1, 1)
    assert info == 4


def test_dense_lu_against_numpy():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((30, 30))
    b = rng.standard_normal((30, 2))
    a, ip, info = algebra.prepare_mat(A)
    assert info == 0
    assert np.abs(algebra.solve_mat(a, ip, b) - np.linalg.solve(A, b)).max() < 1e-11


def test_dense_to_band_rejects_out_of_band():
    A = np.eye(6)
    A[0, 4] = 1.0
    with pytest.raises(ValueError):
        algebra.dense_to_band_storage(A, 1, 1)
