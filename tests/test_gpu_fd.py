"""Config 3 on the GPU: batched LINPACK band LU and one-launch solves of the
finite-difference S, Z and W operators of every degree.

Pivots must equal the oracle's dgbfa restatement bit for bit on the same band
matrices and solutions agree to 1e-12 (north-star bar for config 3); at the
full size (n_r = 512, l_max = 255) the residual of every solve is checked
against the operator itself."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rhs(L, n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((L, n)) + 1j * rng.standard_normal((L, n))


@pytest.mark.parametrize("ob", [2, 3])
def test_fd_band_factor_and_solve_match_oracle(ob):
    import torch
    from magic_amd.fd import FDOperators
    from oracle.algebra import prepare_band, solve_band
    n, l_max = 128, 31
    ops = FDOperators(n, l_max, 6e-5, fd_order=4, fd_order_bound=ob).factor()
    lidx = np.repeat(np.arange(l_max + 1), np.arange(1, l_max + 2)).astype(np.int32)
    L = len(lidx)
    rhs = _rhs(L, n, ob)
    d_l = torch.as_tensor(lidx).cuda()
    d_rhs = torch.as_tensor(rhs).cuda()
    for name in ("S", "Z", "W"):
        abd, fr, fc = getattr(ops, name)
        s = ops.dev[name]
        k = s["k"]
        assert int(s["info"].abs().sum()) == 0
        piv = s["piv"].cpu().numpy()
        lu = s["abd"].cpu().numpy()
        sol = ops.solve(name, d_rhs, d_l).cpu().numpy()
        for l in (0, 1, 2, 17, 31):
            if name != "S" and l == 0:
                continue
            olu, opiv, info = prepare_band(abd[l].copy(), n, k, k)
            assert info == 0
            assert np.array_equal(piv[l], np.asarray(opiv)), (name, l)
            assert np.abs(lu[l] - olu).max() <= 1e-12 * np.abs(olu).max(), (name, l)
            rows = np.nonzero(lidx == l)[0]
            for i in rows[:3]:
                b = rhs[i] * fr[l]
                x = solve_band(olu, n, k, k, opiv, b.real) + 1j * solve_band(olu, n, k, k, opiv, b.imag)
                if name == "W":
                    x = x * fc[l]
                err = np.abs(sol[i] - x).max() / np.abs(x).max()
                assert err < 1e-12, (name, l, i, err)


def test_fd_config3_full_size_residuals():
    """n_r = 512, l_max = 255 (L = 32 896 right-hand sides per system): every
    solve satisfies its (unpreconditioned) banded system to 1e-10 relative."""
    import torch
    from magic_amd.fd import FDOperators
    n, l_max = 512, 255
    ops = FDOperators(n, l_max, 6e-5, fd_order=4, fd_order_bound=2).factor()
    lidx = np.repeat(np.arange(l_max + 1), np.arange(1, l_max + 2)).astype(np.int32)
    L = len(lidx)
    rhs = torch.as_tensor(_rhs(L, n, 5)).cuda()
    d_l = torch.as_tensor(lidx).cuda()
    for name in ("S", "Z", "W"):
        abd, fr, fc = getattr(ops, name)
        k = ops.dev[name]["k"]
        assert int(ops.dev[name]["info"].abs().sum()) == 0
        x = ops.solve(name, rhs, d_l)
        # A_pre = Fr A Fc is stored in abd; A x = b  <=>  A_pre (x / fc) = fr b
        A = torch.as_tensor(abd).cuda()
        frd = torch.as_tensor(fr).cuda()
        fcd = torch.as_tensor(fc).cuda()
        y = x / fcd[lidx.astype(np.int64)]
        Ay = torch.zeros_like(y)
        for dd in range(-k, k + 1):
            lo, hi = max(0, -dd), min(n, n - dd)
            diag = A[:, 2 * k - dd, lo + dd:hi + dd][lidx.astype(np.int64)]
            Ay[:, lo:hi] += diag * y[:, lo + dd:hi + dd]
        b = rhs * frd[lidx.astype(np.int64)]
        rel = float((Ay - b).abs().max() / b.abs().max())
        if name == "S":
            assert rel < 1e-10, (name, rel)
        else:  # degree 0 of Z and W is the identity placeholder: rows l = 0 excluded
            rows = torch.as_tensor(lidx > 0).cuda()
            rel = float((Ay - b)[rows].abs().max() / b[rows].abs().max())
            assert rel < 1e-10, (name, rel)


@pytest.mark.parametrize("ob", [2, 3])
def test_fd_gpu_factors_match_reference_fixture(ob):
    """GPU LINPACK factorisation of every degree against the reference's own
    factored bands and pivots (build_s/z/w_matrices FD branch, fd_order 4,
    n_r_max 64, l_max 12, strat 3, polind 2, g2 1, wimp_lin0 6e-5): pivots
    equal, factors within 1e-13 of the largest entry; a solve with those
    factors agrees with the oracle's dgbsl on the reference factors to 1e-12."""
    import os
    import torch
    from magic_amd.fd import FDOperators
    from oracle.algebra import solve_band
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                             f"fd4_ob{ob}.npz"))
    n, l_max = 64, 12
    ops = FDOperators(n, l_max, 6e-5, fd_order=4, fd_order_bound=ob).factor()
    lidx = np.repeat(np.arange(l_max + 1), np.arange(1, l_max + 2)).astype(np.int32)
    rhs = _rhs(len(lidx), n, 10 + ob)
    d_l, d_rhs = torch.as_tensor(lidx).cuda(), torch.as_tensor(rhs).cuda()
    for name in ("S", "Z", "W"):
        s = ops.dev[name]
        assert int(s["info"].abs().sum()) == 0
        piv, lu = s["piv"].cpu().numpy(), s["abd"].cpu().numpy()
        sol = ops.solve(name, d_rhs, d_l).cpu().numpy()
        k = s["k"]
        for l in range(0 if name == "S" else 1, l_max + 1):
            assert np.array_equal(piv[l], g[f"{name}_piv"][l]), (name, l)
            ref = g[f"{name}_abd"][l]
            assert np.abs(lu[l] - ref).max() <= 1e-13 * np.abs(ref).max(), (name, l)
            fr = g["W_fac_row" if name == "W" else f"{name}_fac"][l]
            i = int(np.nonzero(lidx == l)[0][0])
            b = rhs[i] * fr
            x = (solve_band(ref, n, k, k, g[f"{name}_piv"][l], b.real)
                 + 1j * solve_band(ref, n, k, k, g[f"{name}_piv"][l], b.imag))
            if name == "W":
                x = x * g["W_fac_col"][l]
            assert np.abs(sol[i] - x).max() <= 1e-12 * np.abs(x).max(), (name, l)
