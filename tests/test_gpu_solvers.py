"""GPU parity of the per-degree and banded solver entry points against the oracle.

Banded cases use config-3 sizes (n=512 radial points, l_max=255 -> 256 degrees,
32896 complex right-hand sides, kl=ku=2 and 3 i.e. 5- to 9-wide LINPACK
storage), with seeded diagonally dominant and general band matrices.  Pivots
must equal the oracle's LINPACK sequence exactly; solutions within 1e-12.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _band(n, kl, ku, rng, dominant):
    A = np.zeros((n, n))
    for j in range(n):
        lo, hi = max(0, j - ku), min(n, j + kl + 1)
        A[lo:hi, j] = rng.standard_normal(hi - lo)
    if dominant:
        A[np.arange(n), np.arange(n)] = np.abs(A).sum(axis=1) + 1.0
    else:
        A += 0.5 * np.eye(n)
    return A


def test_chunked_solve_complex_matches_einsum():
    import torch
    from magic_amd.algebra import chunked_solve_complex
    rng = np.random.default_rng(0)
    l_max, n = 20, 33
    lidx = np.concatenate([np.arange(m, l_max + 1) for m in range(l_max + 1)])
    inv = rng.standard_normal((l_max + 1, n, n))
    rhs = rng.standard_normal((lidx.size, n)) + 1j * rng.standard_normal((lidx.size, n))
    got = chunked_solve_complex(torch.from_numpy(inv), torch.from_numpy(lidx),
                                torch.from_numpy(rhs)).cpu().numpy()
    ref = np.einsum("lij,lj->li", inv[lidx], rhs)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("kl,ku,dominant", [(2, 2, True), (3, 3, True), (2, 2, False), (3, 3, False)])
def test_band_factor_pivots_and_solve_config3_sizes(kl, ku, dominant):
    import torch
    from magic_amd import algebra as A
    from oracle import algebra as O
    rng = np.random.default_rng(kl * 10 + int(dominant))
    n, l_max = 512, 255
    nl = l_max + 1
    dense = [_band(n, kl, ku, rng, dominant) for _ in range(8)]
    abd_all = np.stack([O.dense_to_band_storage(dense[l % 8], kl, ku) for l in range(nl)])
    fac = 1.0 / np.abs(np.stack([dense[l % 8] for l in range(nl)])).max(axis=2)
    # oracle factors for the 8 distinct matrices
    ofac = [O.prepare_band(abd_all[l], n, kl, ku) for l in range(8)]
    gabd, gpiv, ginfo = A.prepare_band_batched(torch.from_numpy(abd_all), n, kl, ku)
    gabd, gpiv, ginfo = gabd.cpu().numpy(), gpiv.cpu().numpy(), ginfo.cpu().numpy()
    assert np.all(ginfo == 0)
    for l in range(nl):
        oa, op, oi = ofac[l % 8]
        assert oi == 0
        assert np.array_equal(gpiv[l], op), l                      # pivot order bit for bit
        assert np.abs(gabd[l] - oa).max() <= 1e-13 * np.abs(oa).max()
    lidx = np.concatenate([np.arange(m, l_max + 1) for m in range(l_max + 1)])
    L = lidx.size
    assert L == 32896
    rhs = rng.standard_normal((L, n)) + 1j * rng.standard_normal((L, n))
    sol = A.banded_solve_by_l(torch.from_numpy(gabd), torch.from_numpy(gpiv),
                              torch.from_numpy(lidx), torch.from_numpy(rhs), n, kl, ku,
                              fac_row_by_l=torch.from_numpy(fac)).cpu().numpy()
    # oracle on a sample of degrees (every row of those degrees)
    for l in (0, 1, 7, 100, 255):
        rows = np.nonzero(lidx == l)[0]
        oa, op, _ = ofac[l % 8]
        ref = O.solve_band(oa, n, kl, ku, op, (fac[l][None, :] * rhs[rows]).T).T
        assert np.abs(sol[rows] - ref).max() <= 1e-12 * np.abs(ref).max(), l
        # the factors are those of A itself, so the solve returns A^-1 (fac_row * rhs)
        res = (dense[l % 8] @ sol[rows].T).T - fac[l][None, :] * rhs[rows]
        assert np.abs(res).max() < 1e-10 * np.abs(fac[l][None, :] * rhs[rows]).max()


def test_band_singular_info_and_single_solve():
    import torch
    from magic_amd import algebra as A
    from oracle import algebra as O
    rng = np.random.default_rng(5)
    n, kl, ku = 40, 2, 2
    M = _band(n, kl, ku, rng, True)
    S = M.copy()
    S[:, 7] = 0.0
    abd, piv, info = A.prepare_band(A.dense_to_band_storage(torch.from_numpy(S), kl, ku), n, kl, ku)
    assert info == O.prepare_band(O.dense_to_band_storage(S, kl, ku), n, kl, ku)[2] == 8
    abd, piv, info = A.prepare_band(A.dense_to_band_storage(torch.from_numpy(M), kl, ku), n, kl, ku)
    assert info == 0
    b = rng.standard_normal((n, 3))
    x = A.solve_band_real(abd, n, kl, ku, piv, torch.from_numpy(b)).cpu().numpy()
    assert np.abs(M @ x - b).max() < 1e-11
    with pytest.raises(ValueError):
        A.dense_to_band_storage(torch.eye(6) + torch.diag(torch.ones(2), 4), 1, 1)
