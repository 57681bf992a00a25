"""CPU tests of the product's host-side set-up and of the C ABI library."""

import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from magic_amd import grid
from magic_amd.params import Params
from oracle.params import Params as OP
from oracle import horizontal, radial as oradial

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("l_max,n_phi,L", [(16, 48, 153), (32, 96, 561), (64, 192, 2145),
                                           (85, 256, 3741), (127, 384, 8256), (383, 1152, 73920)])
def test_truncation_sizes(l_max, n_phi, L):
    p = Params({"l_max": l_max})
    assert (p.n_phi, p.n_theta, p.lm_max) == (n_phi, n_phi // 2, L)


@pytest.mark.parametrize("l_max,minc", [(16, 1), (16, 2), (12, 3), (32, 4), (63, 2)])
def test_minc_truncation_and_lm_maps_match_oracle(l_max, minc):
    """Sizes (params.py:60-76) and the m-major lm order of the orders m = 0,
    minc, ...: the same as the oracle's maps, which the reference trajectories
    with minc = 2 and 4 pin; each degree's rows are its orders m <= l."""
    from magic_amd import grid
    from oracle import horizontal
    from oracle.params import Params as OParams
    p, o = Params({"l_max": l_max, "minc": minc}), OParams({"l_max": l_max, "minc": minc})
    assert (p.n_phi, p.n_theta, p.lm_max, p.n_m) == (o.n_phi_max, o.n_theta_max, o.lm_max,
                                                    o.n_m_max)
    assert p.n_phi * minc == p.n_phi_tot
    maps, om = grid.lm_maps(p), horizontal.lm_maps(o)
    assert np.array_equal(maps["l"], om["lm2l"]) and np.array_equal(maps["m"], om["lm2m"])
    for l in range(l_max + 1):
        rows = maps["l_rows"][maps["l_rows_off"][l]:maps["l_rows_off"][l + 1]]
        assert np.array_equal(rows, [om["lm2"][l, m] for m in range(0, l + 1, minc)])
    lay = grid.sht_layout(p, maps)
    K = lay["mp_info"].reshape(p.n_m, 2, 4)[:, :, 0]
    assert K.sum() == p.lm_max   # every (l, m) once, by parity


def test_minc_rejects_sectors_of_odd_length():
    with pytest.raises(ValueError):
        Params({"l_max": 16, "minc": 5})   # n_phi_tot = 48


def test_unknown_key_rejected():
    with pytest.raises(KeyError):
        Params({"l_maxx": 3})


def test_yaml_string_values_and_dt_alias():
    p = Params({"l_max": "24", "ra": "1e5", "l_cour_alf_damp": "false", "dt": 2e-4})
    assert p.l_max == 24 and p.ra == 1e5 and p.l_cour_alf_damp is False and p.dtmax == 2e-4


@pytest.mark.parametrize("n", [24, 48, 96, 576])
def test_gauss_legendre_matches_reference_newton(n):
    x, w = grid.gauss_legendre(n)
    th, wo = horizontal.gauleg(n)
    assert np.abs(np.arccos(x) - th).max() < 1e-13
    assert np.abs(w - wo).max() < 1e-14


def test_radial_operators_match_oracle():
    cfg = {"l_max": 16, "n_r_max": 33, "n_cheb_max": 31}
    rad = grid.radial(Params(cfg))
    op = OP(cfg)
    orad = oradial.radial_functions(op)
    mats = oradial.cheb_mats(op, orad["x"], orad["r_cmb"], orad["r_icb"])
    D1, D2, D3 = oradial.diff_mats(op, mats)
    assert np.abs(rad["r"] - orad["r"]).max() < 1e-15
    for a, b in ((rad["T"], mats[0]), (rad["T1"], mats[1]), (rad["T2"], mats[2]), (rad["T3"], mats[3])):
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
    for a, b in ((rad["D1"], D1), (rad["D2"], D2), (rad["D3"], D3)):
        assert np.abs(a - b).max() <= 1e-11 * np.abs(b).max()
    assert np.abs(rad["C"] - oradial.costf_matrix(33)).max() < 1e-14
    q = oradial.rint_weights(op, orad["r_cmb"], orad["r_icb"])
    assert np.abs(rad["q"] - q).max() < 1e-14


def test_lm_coefficients_match_oracle():
    p = Params({"l_max": 20})
    maps = grid.lm_maps(p)
    c = grid.lm_coefficients(p, maps)
    oc = horizontal.lm_arrays(OP({"l_max": 20}))
    for k in ("dTheta2A", "dTheta2S", "dTheta3A", "dTheta3S", "dTheta4A", "dTheta4S"):
        assert np.abs(c["dTh" + k[6:]] - oc[k]).max() < 1e-12
    om = horizontal.lm_maps(OP({"l_max": 20}))
    assert np.array_equal(maps["l"], om["lm2l"]) and np.array_equal(maps["m"], om["lm2m"])
    # rows of degree l are lm(l, m) for m = 0..l
    for l in (0, 5, 20):
        rows = maps["l_rows"][maps["l_rows_off"][l]:maps["l_rows_off"][l + 1]]
        assert np.all(maps["l"][rows] == l) and np.array_equal(maps["m"][rows], np.arange(l + 1))


@pytest.mark.parametrize("l_max", [33, 383])
def test_sht_layout_covers_every_mode(l_max):
    """Every (m, parity) is covered once by analysis tiles of 16 * waves degrees
    (waves 6 and 4), in 8 equal-length lists (list x: m = x mod 8), padded
    with m = n_m."""
    p = Params({"l_max": l_max})
    lay = grid.sht_layout(p, grid.lm_maps(p))
    info = lay["mp_info"].reshape(p.n_m, 2, 4)
    assert info[:, :, 0].sum() == p.lm_max
    for waves in (6, 4):
        tiles = lay[f"ana_tiles_{waves}"].reshape(8, -1, 3)
        assert tiles.shape[1] * 8 == lay[f"n_ana_tiles_{waves}"]
        for x in range(8):
            real = tiles[x][tiles[x][:, 0] < p.n_m]
            assert np.all(real[:, 0] % 8 == x)
        flat = tiles.reshape(-1, 3)
        for m in range(p.n_m):
            for par in range(2):
                K = info[m, par, 0]
                ks = sorted(t[2] for t in flat if t[0] == m and t[1] == par)
                assert ks == list(range(0, K, 16 * waves))
    assert lay["Hp"] % 32 == 0 and lay["Hp"] >= lay["H"]


@pytest.mark.parametrize("n", [24, 48, 96, 128, 192, 576])
def test_fft_radices(n):
    assert math.prod(grid.fft_radices(n)) == n


def test_noise_matches_oracle():
    from magic_amd.dynamo import seeded_noise
    from oracle.dynamo import seeded_noise as onoise

    class R:
        pass
    p = Params({"l_max": 12, "n_r_max": 17})
    r = R()
    r.maps = grid.lm_maps(p)
    r.rad = grid.radial(p)
    op = OP({"l_max": 12, "n_r_max": 17})
    a = seeded_noise(p, r, 7, 1e-5)
    b = onoise(op, horizontal.lm_maps(op), oradial.radial_functions(op), 7, 1e-5)
    assert np.abs(a - b).max() < 1e-20


def test_conduction_state_matches_oracle():
    from magic_amd.dynamo import conduction_state
    from oracle import matrices

    class R:
        pass
    cfg = {"l_max": 16, "n_r_max": 33}
    p = Params(cfg)
    r = R()
    r.rad = grid.radial(p)
    s0, p0 = conduction_state(p, r)
    op = OP(cfg)
    orad = oradial.radial_functions(op)
    os0, op0 = matrices.ps_cond(op, orad, oradial.cheb_mats(op))
    assert np.abs(s0 - os0).max() < 1e-12 and np.abs(p0 - op0).max() < 1e-9 * np.abs(op0).max()


def test_library_exports_header_symbols():
    from magic_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    hdr = open(os.path.join(ROOT, "include", "magic_hip.h")).read()
    decl = set(re.findall(r"^int (mh_\w+)\(", hdr, flags=re.M))
    assert decl == set(_lib.EXPORTED)
    lib = C.CDLL(_lib.LIB_PATH)
    for name in decl:
        assert hasattr(lib, name), name
    _lib.load()   # struct sizes agree with the binding


def test_noalloc_inverse_checks_sizes_before_touching_a_device():
    """mh_gj_inverse_noalloc refuses a batch beyond its fixed work space (and bad
    sizes) on the host, before any HIP call, so this runs without a GPU."""
    from magic_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()
    n = 300
    fit = _lib.GJ_SCRATCH_DOUBLES // (2 * n * n + 2 * n)
    assert lib.mh_gj_inverse_noalloc(None, None, n, fit + 1, 1, None, None, None) == -3
    assert lib.mh_gj_inverse_noalloc(None, None, 0, 1, 1, None, None, None) == -1
    assert lib.mh_gj_inverse_noalloc(None, None, n, 0, 1, None, None, None) == -1
    with pytest.raises(_lib.MagicHipError, match="work space"):
        _lib.call("mh_gj_inverse_noalloc", None, None, n, fit + 1, 1, None, None, None)


def test_grid_step_rejects_offsets_beyond_32_bits():
    """mh_grid_step forms buffer offsets in 32 bits with 24-bit multiplies and
    refuses, before any launch, shapes whose per-order stride (T*4N*2 elements)
    reaches 2^24 or whose component reaches 2^31 elements; invalid level blocks
    and precisions are refused the same way."""
    from magic_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = _lib.load()

    def call(**kw):
        a = _lib.GridArgs()
        a.N, a.T, a.M, a.Mh = 97, 576, 384, 576
        for k, v in kw.items():
            setattr(a, k, v)
        return lib.mh_grid_step(C.byref(a), None)

    assert call(N=2731) == -1          # 576 * 4 * 2731 * 2 = 12.58e6 < 2^24 but M*... >= 2^31
    assert call(T=1536, N=1366) == -1  # 1536 * 4 * 1366 * 2 = 16.79e6 >= 2^24
    assert call(rb=99) == -1
    assert call(fprec=2) == -1


def test_solver_entry_points_check_shapes_before_launch():
    """The drop-in solvers refuse shapes and degree indices that would make a
    kernel read past its buffers (checked on the host, before any device work:
    these raise on a machine without a GPU too)."""
    import numpy as np
    from magic_amd import algebra
    n, kl, ku, nl = 16, 2, 2, 4
    abd = np.zeros((nl, 2 * kl + ku + 1, n))
    piv = np.ones((nl, n), np.int64)
    rhs = np.zeros((6, n), np.complex128)
    lidx = np.array([0, 1, 1, 2, 3, 3])
    with pytest.raises(IndexError):
        algebra.banded_solve_by_l(abd, piv, np.array([0, 1, 1, 2, 3, 4]), rhs, n, kl, ku)
    with pytest.raises(IndexError):
        algebra.banded_solve_by_l(abd, piv, np.array([0, 1, 1, 2, 3, -1]), rhs, n, kl, ku)
    with pytest.raises(ValueError):
        algebra.banded_solve_by_l(abd, piv, lidx[:5], rhs, n, kl, ku)
    with pytest.raises(ValueError):
        algebra.banded_solve_by_l(abd, piv, lidx, rhs[:, :n - 1], n, kl, ku)
    with pytest.raises(ValueError):
        algebra.banded_solve_by_l(abd[:, :6], piv, lidx, rhs, n, kl, ku)
    with pytest.raises(ValueError):
        algebra.banded_solve_by_l(abd, piv[:3], lidx, rhs, n, kl, ku)
    with pytest.raises(ValueError):
        algebra.banded_solve_by_l(abd, piv, lidx, rhs, n, kl, ku, fac_row_by_l=np.ones((nl, 3)))
    with pytest.raises(ValueError):
        algebra.banded_solve_by_l(np.zeros((nl, 16, n)), piv, lidx, rhs, n, 5, 5)
    with pytest.raises(ValueError):
        algebra.solve_band_real(abd[0], n, kl, ku, piv[0], np.zeros(n + 1))
    with pytest.raises(ValueError):
        algebra.prepare_band(np.zeros((6, n)), n, kl, ku)
    inv = np.zeros((nl, n, n))
    with pytest.raises(IndexError):
        algebra.chunked_solve_complex(inv, np.array([0, 1, 1, 2, 3, 9]), rhs)
    with pytest.raises(ValueError):
        algebra.chunked_solve_complex(inv, lidx, rhs[:, :3])
    with pytest.raises(ValueError):
        algebra.chunked_solve_complex(inv[:, :, :3], lidx, rhs)


def test_bench_defines_every_function_main_calls():
    """bench.py's main() calls module functions by name; a missing one would
    only fail on the GPU box at the end of a run."""
    import builtins
    import dis
    import bench
    called = {ins.argval for ins in dis.get_instructions(bench.main)
              if ins.opname in ("LOAD_GLOBAL", "LOAD_NAME")}
    missing = [n for n in called if not hasattr(bench, n) and not hasattr(builtins, n)]
    assert not missing, missing


def test_inner_core_grid_and_coupled_operators_match_oracle():
    """Conducting inner core: the host inner grid, even-Chebyshev polynomials
    and derivative operators, and the coupled B/J inverses and row operators,
    against the oracle (pinned by the reference's conducting-core trajectories):
    applying the row operators to a right-hand side gives the oracle's solve ->
    DCT-I -> derivatives of both shells."""
    from magic_amd import grid, inner_core
    from oracle import matrices, radial as orad
    from oracle.dynamo import OracleRun
    cfg = dict(l_max=8, n_r_max=17, n_cheb_max=15, n_cheb_ic_max=7, sigma_ratio=0.7, kbotb=3)
    p, o = Params(cfg), OracleRun(cfg)
    op = o.prm
    assert (p.n_r_ic_max, p.n_r_tot, p.n_cheb_ic_max) == (op.n_r_ic_max, op.n_r_tot, 7)
    rad = grid.radial(p)
    ic = grid.inner_core(p, rad["r_icb"])
    oic = orad.ic_grid(op, o.rad["r_icb"])
    for k in ("r_ic", "O_r_ic", "cheb", "dcheb", "d2cheb"):
        np.testing.assert_allclose(ic[k], oic[k], rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(ic["D1"], oic["D1"].T, rtol=1e-10, atol=1e-10)
    w0 = 3e-4
    bi, ji = inner_core.coupled_inverses(p, rad, ic, w0)
    obi, oji = matrices.build_bj_coupled(op, o.rad, o.mats, oic, w0)
    for a, b in ((bi, obi), (ji, oji)):
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
    N, Nic, NT = p.n_r_max, p.n_r_ic_max, p.n_r_tot
    k4 = -(-NT // 4) * 4
    (mo, Ro, _), (mi, Ri, _) = inner_core.row_operators(p, rad, ic, bi, k4)
    rng = np.random.default_rng(5)
    l = 3
    rhs = rng.standard_normal(NT)
    sol = obi[l] @ rhs
    c_oc, c_ic = sol[:N].copy(), sol[N:].copy()
    c_oc[op.n_cheb_max:] = 0.0
    c_ic[op.n_cheb_ic_max:] = 0.0
    f_oc, f_ic = orad.costf(c_oc), orad.costf(c_ic)
    ref_oc = np.concatenate([f_oc, o.D1 @ f_oc, o.D2 @ f_oc])
    ref_ic = np.concatenate([f_ic, f_ic @ oic["D1"], f_ic @ oic["D2"]])
    got_oc = mo[l, :NT, :Ro].T @ rhs
    got_ic = mi[l, :NT, :Ri].T @ rhs
    assert np.abs(got_oc - ref_oc).max() <= 1e-9 * np.abs(ref_oc).max()
    assert np.abs(got_ic - ref_ic).max() <= 1e-9 * np.abs(ref_ic).max()
    p.check_step_supported()   # conducting: on the device path
    Params(dict(cfg, nRotIC=1)).check_step_supported()   # and rotating (no-slip ICB)
    for bad in (dict(nRotIC=-1), dict(nRotIC=1, sigma_ratio=0.0), dict(nRotIC=1, kbotv=1),
                dict(nRotIC=1, mode=1)):
        with pytest.raises(NotImplementedError):
            Params(dict(cfg, **bad)).check_step_supported()


def test_rotating_inner_core_operator_matches_oracle():
    """Rotating conducting inner core: the rotation constants and the z(l=1, m=0)
    inverse against the oracle (pinned by the reference's rotating-core
    trajectories, omega_ic included), and its row operator applied to a
    right-hand side against the oracle's solve -> truncation -> DCT-I ->
    derivatives."""
    from magic_amd import inner_core
    from oracle import matrices, radial as orad
    from oracle.dynamo import OracleRun
    cfg = dict(l_max=8, n_r_max=17, n_cheb_max=15, n_cheb_ic_max=7, sigma_ratio=0.7, kbotb=3,
               nRotIC=1, minc=2)
    p, o = Params(cfg), OracleRun(cfg)
    c = inner_core.rotation_constants(p)
    assert math.isclose(c["c_z10_omega"], o.c_z10_omega_ic, rel_tol=1e-15)
    assert math.isclose(c["c_dt_z10"], o.c_dt_z10_ic, rel_tol=1e-15)
    assert math.isclose(c["c_lorentz"], o.c_lorentz_ic, rel_tol=1e-15)
    assert math.isclose(c["fac_lt"], o.prm.LFfac * 2.0 * math.pi / o.prm.n_phi_max, rel_tol=1e-15)
    rad = grid.radial(p)
    w0 = 2.5e-4
    inv = inner_core.z10_inverse(p, rad, w0, c["c_dt_z10"])
    oinv = matrices.build_z10(o.prm, o.rad, o.mats, w0, o.c_dt_z10_ic)
    assert np.abs(inv - oinv).max() <= 1e-10 * np.abs(oinv).max()
    N = p.n_r_max
    k4, rpad = -(-N // 4) * 4, -(-(3 * N) // 64) * 64
    mt = inner_core.z10_row_operator(p, rad, inv, k4, rpad)
    assert mt.shape == (k4, rpad) and not mt[N:].any() and not mt[:, 3 * N:].any()
    rhs = np.random.default_rng(7).standard_normal(N)
    sol = oinv @ rhs
    sol[p.n_cheb_max:] = 0.0
    f = orad.costf(sol)
    ref = np.concatenate([f, o.D1 @ f, o.D2 @ f])
    got = mt[:N, :3 * N].T @ rhs
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


@pytest.mark.parametrize("cfg", [
    dict(l_max=12, n_r_max=17, n_cheb_max=15, mode=1, strat=5.0, polind=2.0, g2=1.0,
         ktopv=1, kbotv=1, ra=1.48638035e5),
    dict(l_max=12, n_r_max=17, strat=1.0, polind=2.0, g1=1.0, ra=1e5, prmag=5.0)])
def test_anelastic_operators_match_oracle(cfg):
    """Anelastic host side (magic_amd.anelastic) against the oracle (pinned by
    the reference's anelastic trajectories): background profiles and factors,
    the S, Z, (w, p) and l = 0 pressure inverses, the conduction state, the
    angular-momentum factors, and the row operators applied to a right-hand
    side against the oracle's solve -> truncation -> DCT-I -> derivatives."""
    from magic_amd import anelastic
    from oracle import anelastic as oan, radial as orad
    from oracle.dynamo import OracleRun
    p, o = Params(dict(cfg, l_correct_AMz=True)), OracleRun(dict(cfg, l_correct_AMz=True))
    op = o.prm
    rad = grid.radial(p)
    prof = anelastic.background(p, rad)
    for k in ("rgrav", "temp0", "rho0", "orho1", "otemp1", "beta", "dbeta", "dLtemp0"):
        np.testing.assert_allclose(prof[k], o.prof[k], rtol=1e-13, atol=1e-13)
    for k in ("ViscHeatFac", "OhmLossFac"):
        assert prof[k] == pytest.approx(o.prof[k], rel=1e-14)
    w0 = 3e-4
    pairs = ((anelastic.s_inverse(p, rad, prof, w0), oan.build_s(op, o.rad, o.mats, o.prof, w0)),
             (anelastic.z_inverse(p, rad, prof, w0), oan.build_z(op, o.rad, o.mats, o.prof, w0)),
             (anelastic.wp_inverse(p, rad, prof, w0), oan.build_wp(op, o.rad, o.mats, o.prof, w0)),
             (anelastic.p0_inverse(p, rad, prof)[None], oan.build_p0(op, o.rad, o.mats, o.prof)[None]))
    for a, b in pairs:
        for l in range(a.shape[0]):
            assert np.abs(a[l] - b[l]).max() <= 1e-8 * max(np.abs(b[l]).max(), 1e-300), l
    assert anelastic.p0_integral(p, prof) == oan.p0_integral_bc(op, o.prof)
    np.testing.assert_allclose(anelastic.p0_weights(p, rad, prof),
                               oan.p0_rhs_weights(op, o.rad, o.prof, o.q), rtol=1e-12, atol=1e-300)
    s0, p0 = anelastic.conduction_state(p, rad, prof)
    os0, op0 = oan.ps_cond(op, o.rad, o.mats, o.prof)
    assert np.abs(s0 - os0).max() <= 1e-11 * np.abs(os0).max()
    assert np.abs(p0 - op0).max() <= 1e-8 * np.abs(op0).max()
    am, oam = anelastic.am_setup(p, rad, prof), o.am
    for k, ok in (("r2q", "r2q"), ("pz", "z"), ("pdz", "dz"), ("pd2z", "d2z")):
        np.testing.assert_allclose(am[k], oam[ok], rtol=1e-12, atol=1e-300)
    assert am["cz"] == pytest.approx(8.0 / 3.0 * math.pi / oam["c_moi_oc"], rel=1e-13)
    # the device rebuild's operators: A0 + w0 A1 is the collocation matrix (its
    # inverse is the oracle's), and the l = 0 pressure block maps the pressure half
    # of a right-hand side to (p, dp/dr)
    N = p.n_r_max
    ops = anelastic.affine_operators(p, rad, prof)
    (a0, a1), (w0m, w1m) = ops["sz"], ops["wp"]
    for k, inv in ((0, pairs[0][1]), (1, pairs[1][1])):
        for l in range(1, p.l_max + 1):
            M = np.linalg.inv(a0[k, l] + w0 * a1[k, l])
            assert np.abs(M - inv[l]).max() <= 1e-8 * np.abs(inv[l]).max(), (k, l)
    for l in range(1, p.l_max + 1):
        A = w0m[l] + w0 * w1m[l]
        x = np.random.default_rng(l).standard_normal(2 * N)
        y = pairs[2][1][l] @ x
        assert np.abs(A @ y - x).max() <= 1e-7 * np.abs(x).max(), l
    k4w, r6 = -(-(2 * N) // 4) * 4, -(-(6 * N) // 64) * 64
    mt0 = anelastic.p0_block(p, rad, prof, k4w, r6)
    rhs = np.random.default_rng(11).standard_normal(N)
    cp = pairs[3][1][0] @ rhs
    cp[p.n_cheb_max:] = 0.0
    fp = orad.costf(cp)
    got = mt0[N:2 * N, 4 * N:6 * N].T @ rhs
    ref = np.concatenate([fp, o.D1 @ fp])
    assert np.abs(got - ref).max() <= 1e-8 * np.abs(ref).max()
    assert not np.abs(mt0[:N]).any()


@pytest.mark.parametrize("cfg", [
    dict(l_max=8, n_r_max=33, radial_scheme="FD", fd_order=2, fd_order_bound=2, prmag=5.0),
    dict(l_max=8, n_r_max=40, radial_scheme="FD", fd_order=4, fd_order_bound=2, prmag=5.0,
         ktopv=1, kbotv=1),
    dict(l_max=8, n_r_max=33, radial_scheme="FD", strat=0.5, polind=2.0, g0=0.0, g1=0.0,
         g2=1.0, mode=1, ktopv=1, kbotv=1)])
def test_finite_difference_host_operators_match_oracle(cfg):
    """Finite-difference set-up of the device step against the oracle (itself
    pinned by reference FD trajectories): grid, dense stencil matrices, Simpson
    weights, inverses of A0 + w A1 for S, Z, W (double curl), B, J, the l = 0
    pressure inverse and the conduction state."""
    from magic_amd import anelastic, fdstep
    from oracle import anelastic as oan, fdstep as ofd, radial as orad
    from oracle.params import Params as OP
    op, p = OP(cfg), Params(cfg)
    orr = orad.radial_functions(op)
    fm = orad.fd_mats(op, orr["r"])
    oprof = oan.background(op, orr)
    rad = fdstep.radial(p)
    assert np.array_equal(rad["r"], orr["r"])
    for a, b in zip((rad["D1"], rad["D2"], rad["D3"], rad["D4"]), fm[1:]):
        assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()
    assert np.abs(rad["q"] - orad.simps_weights(orr["r"])).max() < 1e-15
    prof = anelastic.background(p, rad) if p.l_anel else fdstep.boussinesq_profile(rad)
    w0 = 3e-5
    inv = ofd.build(op, orr, fm, oprof, w0)
    ops = fdstep.affine_operators(p, rad, prof)
    (A0, A1), (W0, W1) = ops["szbj"], ops["w"]
    for k, name in enumerate(("s", "z", "b", "j")[:A0.shape[0]]):
        for l in range(0 if name == "s" else 1, p.l_max + 1):
            M = np.linalg.inv(A0[k, l] + w0 * A1[k, l])
            assert np.abs(M - inv[name][l]).max() <= 1e-10 * np.abs(inv[name][l]).max(), (name, l)
    for l in range(1, p.l_max + 1):
        M = np.linalg.inv(W0[l] + w0 * W1[l])
        assert np.abs(M - inv["w"][l]).max() <= 1e-9 * np.abs(inv["w"][l]).max(), ("w", l)
    assert np.abs(fdstep.p0_inverse(p, rad, prof) - inv["p0"]).max() < 1e-12
    s0, p0 = fdstep.conduction_state(p, rad, prof)
    os0, op0 = ofd.ps_cond(op, orr, fm, oprof)
    assert np.abs(s0 - os0).max() < 1e-12 and np.abs(p0 - op0).max() < 1e-12 * np.abs(op0).max()
