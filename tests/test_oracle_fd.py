"""Pins of the oracle's finite-difference restatement (oracle/fd.py): exactness
on polynomials, grid shape, closed forms of the polytropic state, band widths
and LINPACK pivots of the operators against scipy's dgbtrf."""

import math

import numpy as np
import pytest

from oracle import fd
from oracle.algebra import dense_to_band_storage, prepare_band


def test_fornberg_exact_on_polynomials():
    rng = np.random.default_rng(0)
    x = np.sort(rng.uniform(-1.0, 1.0, 7))
    z = 0.13
    c = fd.fornberg(z, x, 6, 4)
    for p in range(7):
        for k in range(5):
            got = np.dot(c[:, k], x ** p)
            exact = 0.0 if k > p else math.factorial(p) / math.factorial(p - k) * z ** (p - k)
            assert abs(got - exact) < 1e-9 * (1 + abs(exact)), (p, k, got, exact)


def _grid(n=65):
    rcmb = 1.0 / (1.0 - 0.35)
    return fd.fd_grid(n, rcmb - 1.0, rcmb, 0.3, 0.1), rcmb


def test_stretched_grid_shape():
    r, rcmb = _grid(129)
    dr = -np.diff(r)
    assert r[0] == rcmb and r[-1] == rcmb - 1.0
    assert np.all(dr > 0)
    # fine boundary spacing / bulk spacing is the requested ratio 0.1 (to rounding)
    assert abs(dr[0] / dr.max() - 0.1) < 1e-6 and abs(dr[-1] / dr.max() - 0.1) < 1e-2
    ru = fd.fd_grid(11, rcmb - 1.0, rcmb, 0.3, 1.0)
    assert np.allclose(-np.diff(ru), 0.1)


@pytest.mark.parametrize("order,ob", [(2, 2), (4, 2), (4, 3)])
def test_derivative_matrices_exact_on_polynomials(order, ob):
    r, _ = _grid(41)
    mats = fd.fd_der_mats(r, fd.fd_coeffs(r, order, ob), order, ob)
    x = r - r.mean()
    # bulk stencils are exact for degree <= order (+2 for d3/d4), boundary ones for ob (+1, +2, +3)
    for k, deg in ((1, min(order, ob)), (2, min(order, ob + 1)), (3, min(order + 2, ob + 2)),
                   (4, min(order + 2, ob + 3))):
        if order == 2 and k >= 3:
            continue  # order-2 3rd/4th derivatives use ghost points at the walls
        for p in range(deg + 1):
            f = x ** p
            exact = (math.factorial(p) / math.factorial(p - k) * x ** (p - k)) if p >= k else 0 * x
            got = mats[k] @ f
            scale = np.abs(mats[k]) @ np.abs(f)   # size of the cancelling terms
            assert np.all(np.abs(got - exact) <= 1e-12 * scale + 1e-12), (k, p)


def test_polytropic_closed_forms():
    r, rcmb = _grid(257)
    prof = fd.polytropic(r, rcmb, rcmb - 1.0, 0.35, 3.0, 2.0, 0.0, 0.0, 1.0)
    assert np.allclose(prof["rho0"], prof["temp0"] ** 2.0)
    # density contrast exp(strat) between inner and outer boundary
    assert abs(prof["rho0"][-1] / prof["rho0"][0] - math.exp(3.0)) < 1e-10 * math.exp(3.0)
    # beta = d ln rho0 / dr, dbeta = d beta / dr (centred differences of the closed forms)
    h = 1e-6
    for i in (3, 100, 250):
        rp, rm = r[i] + h, r[i] - h
        pp = fd.polytropic(np.array([rp]), rcmb, rcmb - 1.0, 0.35, 3.0, 2.0, 0.0, 0.0, 1.0)
        pm = fd.polytropic(np.array([rm]), rcmb, rcmb - 1.0, 0.35, 3.0, 2.0, 0.0, 0.0, 1.0)
        assert abs((math.log(pp["rho0"][0]) - math.log(pm["rho0"][0])) / (2 * h)
                   - prof["beta"][i]) < 1e-6 * abs(prof["beta"][i])
        assert abs((pp["beta"][0] - pm["beta"][0]) / (2 * h) - prof["dbeta"][i]) < 1e-5 * abs(prof["dbeta"][i])
        assert abs((pp["dbeta"][0] - pm["dbeta"][0]) / (2 * h) - prof["ddbeta"][i]) < 1e-4 * abs(prof["ddbeta"][i])


@pytest.mark.parametrize("ob", [2, 3])
def test_operators_fit_band_and_match_lapack_pivots(ob):
    from scipy.linalg import lapack
    n, order = 96, 4
    r, rcmb = _grid(n)
    mats = fd.fd_der_mats(r, fd.fd_coeffs(r, order, ob), order, ob)
    prof = fd.polytropic(r, rcmb, rcmb - 1.0, 0.35, 3.0, 2.0, 0.0, 0.0, 1.0)
    hs, hw = fd.band_widths(order, ob)
    ops = [(fd.s_operator(7, 6e-5, 1.0, mats, prof)[0], hs),
           (fd.z_operator(7, 6e-5, mats, prof)[0], hs),
           (fd.w_operator(7, 6e-5, mats, prof)[0], hw)]
    for A, k in ops:
        abd = dense_to_band_storage(A, k, k)            # raises if A leaves the band
        lu, piv, info = prepare_band(abd.copy(), n, k, k)
        assert info == 0
        ab = np.zeros((3 * k + 1, n))
        for j in range(n):
            for i in range(max(0, j - k), min(n, j + k + 1)):
                ab[2 * k + i - j, j] = A[i, j]
        _, ipiv, info2 = lapack.dgbtrf(ab, k, k)
        assert info2 == 0
        assert np.array_equal(np.asarray(piv), ipiv + 1)


# ------------------------------------------------------------ reference fixtures
def _fd_fixture(ob):
    import os
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    return np.load(os.path.join(gold, f"fd4_ob{ob}.npz"))


@pytest.mark.parametrize("ob", [2, 3])
def test_operators_match_reference_factored_bands(ob):
    """The reference's own FD build (build_s_matrices, build_z_matrices,
    build_w_matrices at fd_order 4, n_r_max 64, l_max 12, strat 3, polind 2,
    g2 1, wimp_lin0 6e-5; tests/golden/fd4_ob*.npz): grid, polytropic profiles,
    and per degree the LINPACK-factored band, the pivots and the row (and for W
    column) preconditioning factors."""
    g = _fd_fixture(ob)
    n, order = 64, 4
    r, rcmb = _grid(n)
    assert np.abs(r - g["r"]).max() <= 1e-15 * rcmb
    prof = fd.polytropic(r, rcmb, rcmb - 1.0, 0.35, 3.0, 2.0, 0.0, 0.0, 1.0)
    for k in ("rho0", "temp0", "beta", "dbeta", "ddbeta", "dLtemp0", "orho1", "rgrav"):
        assert np.abs(prof[k] - g["prof_" + k]).max() <= 1e-14 * np.abs(g["prof_" + k]).max(), k
    mats = fd.fd_der_mats(r, fd.fd_coeffs(r, order, ob), order, ob)
    hs, hw = fd.band_widths(order, ob)
    assert tuple(g["S_k"]) == tuple(g["Z_k"]) == (hs, hs) and tuple(g["W_k"]) == (hw, hw)
    for l in range(13):
        ops = [("S", fd.s_operator(l, 6e-5, 1.0, mats, prof), hs)]
        if l >= 1:   # the reference keeps identity bands for Z and W at l = 0
            ops += [("Z", fd.z_operator(l, 6e-5, mats, prof), hs),
                    ("W", fd.w_operator(l, 6e-5, mats, prof), hw)]
        for name, op, k in ops:
            lu, piv, info = prepare_band(dense_to_band_storage(op[0], k, k).copy(), n, k, k)
            assert info == 0
            assert np.array_equal(np.asarray(piv), g[f"{name}_piv"][l]), (name, l)
            ref = g[f"{name}_abd"][l]
            assert np.abs(lu - ref).max() <= 1e-13 * np.abs(ref).max(), (name, l)
            fr = g["W_fac_row" if name == "W" else f"{name}_fac"][l]
            assert np.abs(op[1] - fr).max() <= 1e-14 * np.abs(fr).max(), (name, l)
            if name == "W":
                fc = g["W_fac_col"][l]
                assert np.abs(op[2] - fc).max() <= 1e-14 * np.abs(fc).max(), l
