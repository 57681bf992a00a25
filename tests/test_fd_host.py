"""Host-side FD operators (magic_amd.fd, band-native, all degrees at once)
against the oracle's dense per-degree restatement of the reference builders."""

import numpy as np
import pytest

from magic_amd import fd as pfd
from oracle import fd as ofd


def _dense(abd, k, n):
    A = np.zeros((n, n))
    for j in range(n):
        for i in range(max(0, j - k), min(n, j + k + 1)):
            A[i, j] = abd[2 * k + i - j, j]
    return A


def test_grid_and_stencils_match_oracle():
    rcmb = 1.0 / 0.65
    for n, ratio in ((65, 0.1), (129, 0.1), (33, 1.0)):
        rp = pfd.stretched_grid(n, rcmb - 1.0, rcmb, 0.3, ratio)
        ro = ofd.fd_grid(n, rcmb - 1.0, rcmb, 0.3, ratio)
        assert np.abs(rp - ro).max() < 1e-14
    r = ofd.fd_grid(64, rcmb - 1.0, rcmb, 0.3, 0.1)
    for order, ob in ((4, 2), (4, 3), (2, 2)):
        hw = 12
        bands = pfd.derivative_bands(r, order, ob, hw)
        mats = ofd.fd_der_mats(r, ofd.fd_coeffs(r, order, ob), order, ob)
        for k in (1, 2, 3, 4):
            if order == 2 and k >= 3:
                continue
            D = np.zeros((64, 64))
            for i in range(64):
                for dd in range(-hw, hw + 1):
                    if 0 <= i + dd < 64:
                        D[i, i + dd] = bands[k][i, dd + hw]
            scale = np.abs(mats[k]).max()
            assert np.abs(D - mats[k]).max() <= 1e-13 * scale, (order, ob, k)


@pytest.mark.parametrize("ob", [2, 3])
def test_operators_match_oracle(ob):
    n, l_max, wimp = 80, 12, 6e-5
    ops = pfd.FDOperators(n, l_max, wimp, fd_order=4, fd_order_bound=ob)
    rcmb = 1.0 / 0.65
    r = ofd.fd_grid(n, rcmb - 1.0, rcmb, 0.3, 0.1)
    mats = ofd.fd_der_mats(r, ofd.fd_coeffs(r, 4, ob), 4, ob)
    prof = ofd.polytropic(r, rcmb, rcmb - 1.0, 0.35, 3.0, 2.0, 0.0, 0.0, 1.0)
    hs, hw = ofd.band_widths(4, ob)
    assert (ops.kls, ops.klw) == (hs, hw)
    for l in (0, 1, 5, 12):
        A, fr = ofd.s_operator(l, wimp, 1.0, mats, prof)
        assert np.abs(_dense(ops.S[0][l], hs, n) - A).max() < 1e-13
        assert np.abs(ops.S[1][l] - fr).max() < 1e-13 * np.abs(fr).max()
        if l == 0:
            continue
        A, fr = ofd.z_operator(l, wimp, mats, prof)
        assert np.abs(_dense(ops.Z[0][l], hs, n) - A).max() < 1e-13
        A, fr, fc = ofd.w_operator(l, wimp, mats, prof)
        assert np.abs(_dense(ops.W[0][l], hw, n) - A).max() < 1e-12
        assert np.abs(ops.W[2][l] - fc).max() < 1e-12 * np.abs(fc).max()


@pytest.mark.parametrize("ob", [2, 3])
def test_host_bands_factor_to_the_reference_bands(ob):
    """The band-native host build of every degree, factored in LINPACK order
    on the CPU, against the reference's own factored bands, pivots and
    preconditioning factors (tests/golden/fd4_ob*.npz: fd_order 4, n_r_max 64,
    l_max 12, strat 3, polind 2, g2 1, wimp_lin0 6e-5).  The GPU factorisation
    of the same bands is checked against this fixture in test_gpu_fd.py."""
    import os
    from oracle.algebra import prepare_band
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             f"fd4_ob{ob}.npz"))
    ops = pfd.FDOperators(64, 12, 6e-5, fd_order=4, fd_order_bound=ob)
    assert np.abs(ops.r - g["r"]).max() <= 1e-15 * ops.r[0]
    for name, (abd, fr, fc), k in (("S", ops.S, ops.kls), ("Z", ops.Z, ops.kls),
                                   ("W", ops.W, ops.klw)):
        assert tuple(g[f"{name}_k"]) == (k, k)
        for l in range(0 if name == "S" else 1, 13):
            lu, piv, info = prepare_band(abd[l].copy(), 64, k, k)
            assert info == 0
            assert np.array_equal(np.asarray(piv), g[f"{name}_piv"][l]), (name, l)
            ref = g[f"{name}_abd"][l]
            assert np.abs(lu - ref).max() <= 1e-13 * np.abs(ref).max(), (name, l)
            rf = g["W_fac_row" if name == "W" else f"{name}_fac"][l]
            assert np.abs(fr[l] - rf).max() <= 1e-14 * np.abs(rf).max(), (name, l)
            if name == "W":
                assert np.abs(fc[l] - g["W_fac_col"][l]).max() <= 1e-14 * np.abs(fc[l]).max()
