"""GPU parity of the individual HIP kernels against the CPU oracle / numpy.

Each test calls the kernels through the C ABI (ctypes binding) on torch device
buffers and compares with the FP64 oracle restatement of the reference.
"""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def test_gemm_f64_matches_numpy():
    """FP64 MFMA GEMM (also checks the 16x16x4 lane layout) against numpy."""
    torch = _torch()
    from magic_amd import _lib
    from magic_amd.device import ptr, stream_handle
    rng = np.random.default_rng(1)
    for (M, Nn, K, batch) in ((16, 64, 4, 1), (37, 91, 53, 3), (97, 291, 97, 2)):
        A = rng.standard_normal((batch, M, K))
        B = rng.standard_normal((batch, K, Nn))
        dA = torch.from_numpy(A).cuda()
        dB = torch.from_numpy(B).cuda()
        dC = torch.zeros(batch, M, Nn, dtype=torch.float64, device="cuda")
        g = _lib.GemmArgs()
        g.A, g.B, g.C = ptr(dA), ptr(dB), ptr(dC)
        g.sAr, g.sAc, g.sAb = K, 1, M * K
        g.sBr, g.sBc, g.sBb = Nn, 1, K * Nn
        g.sCr, g.sCc, g.sCb = Nn, 1, M * Nn
        g.M, g.Nn, g.K, g.batch, g.alpha = M, Nn, K, batch, 1.0
        _lib.call("mh_gemm_f64", _lib.ref(g), stream_handle())
        ref = A @ B
        err = np.abs(dC.cpu().numpy() - ref).max() / np.abs(ref).max()
        assert err < 1e-14, (M, Nn, K, err)


@pytest.mark.parametrize("l_max,minc", [(16, 1), (32, 1), (63, 1), (16, 2), (32, 4)])
def test_plm_tables_match_oracle(l_max, minc):
    """Device P_lm / dP_lm tables against the oracle's recurrence (plms.py);
    minc > 1: the tables of the orders m = 0, minc, ... in the run's lm order."""
    torch = _torch()
    from magic_amd.params import Params
    from magic_amd.device import Resolution
    from oracle.params import Params as OP
    from oracle import horizontal
    p = Params({"l_max": l_max, "n_r_max": 9, "minc": minc})
    res = Resolution(p)
    op = OP({"l_max": l_max, "n_r_max": 9, "minc": minc})
    tab = horizontal.plm_tables(op)
    lay = res.lay
    info = lay["mp_info"].reshape(p.n_m, 2, 4)
    synP = res.synP.cpu().numpy()
    synD = res.synD.cpu().numpy()
    anaW = res.anaW.cpu().numpy()
    H, Hp = lay["H"], lay["Hp"]
    lm = 0
    worst = 0.0
    for mi in range(p.n_m):
        m = mi * minc
        for l in range(m, l_max + 1):
            par, k = (l - m) % 2, (l - m) // 2
            K, so, ao, K16 = info[mi, par]
            gP = synP[so + k * Hp: so + k * Hp + H]
            gD = synD[so + k * Hp: so + k * Hp + H]
            gW = anaW[ao + np.arange(H) * K16 + k]
            scale = max(np.abs(tab["Plm"][lm]).max(), 1e-300)
            worst = max(worst, np.abs(gP - tab["Plm"][lm]).max() / scale,
                        np.abs(gD - tab["dPlm"][lm]).max() / max(np.abs(tab["dPlm"][lm]).max(), 1e-300),
                        np.abs(gW - tab["wPlm"][lm]).max() / max(np.abs(tab["wPlm"][lm]).max(), 1e-300))
            lm += 1
    assert lm == p.lm_max
    assert worst < 1e-12, worst


def _oracle_sht(l_max, minc=1):
    from oracle.params import Params as OP
    from oracle.sht import SHT as OSHT
    return OSHT(OP({"l_max": l_max, "n_r_max": 9, "minc": minc}))


def _rand_spec(L, B, lmaps, seed):
    rng = np.random.default_rng(seed)
    l = lmaps["l"]
    x = (rng.standard_normal((L, B)) + 1j * rng.standard_normal((L, B))) / (l[:, None] + 1.0) ** 2
    x[lmaps["m"] == 0] = x[lmaps["m"] == 0].real
    return x


@pytest.mark.parametrize("l_max,B,minc", [(16, 1, 1), (16, 7, 1), (32, 33, 1), (64, 20, 1),
                                          (16, 7, 2), (12, 5, 3), (32, 9, 4)])
def test_sht_matches_oracle(l_max, B, minc):
    """scal_to_spat / scal_to_SH / torpol_to_spat / spat_to_sphertor vs the
    oracle; minc > 1: the orders m = 0, minc, ... on one 2 pi / minc sector."""
    torch = _torch()
    from magic_amd.params import Params
    from magic_amd.device import Resolution
    from magic_amd.sht import SHT
    p = Params({"l_max": l_max, "n_r_max": 9, "minc": minc})
    res = Resolution(p)
    sht = SHT(res)
    osht = _oracle_sht(l_max, minc)
    L = p.lm_max
    S = _rand_spec(L, B, res.maps, 1)
    Q = _rand_spec(L, B, res.maps, 2)
    T = _rand_spec(L, B, res.maps, 3)
    dev = lambda a: torch.from_numpy(a).cuda()
    # scalar synthesis
    g = sht.scal_to_spat(dev(S)).cpu().numpy()
    go = osht.scal_to_spat(S)
    assert np.abs(g - go).max() / np.abs(go).max() < 1e-13
    # scalar analysis of the oracle grid
    s2 = sht.scal_to_SH(dev(go)).cpu().numpy()
    so = osht.scal_to_SH(go)
    assert np.abs(s2 - so).max() / np.abs(so).max() < 1e-13
    assert np.abs(s2 - S).max() / np.abs(S).max() < 1e-12   # round trip
    # vector synthesis
    br, bt, bp = (x.cpu().numpy() for x in sht.torpol_to_spat(dev(Q), dev(S), dev(T)))
    obr, obt, obp = osht.torpol_to_spat(Q, S, T)
    for a, b in ((br, obr), (bt, obt), (bp, obp)):
        assert np.abs(a - b).max() / np.abs(b).max() < 1e-13
    # vector analysis
    Sv, Tv = (x.cpu().numpy() for x in sht.spat_to_sphertor(dev(obt), dev(obp)))
    oS, oT = osht.spat_to_sphertor(obt, obp)
    assert np.abs(Sv - oS).max() / np.abs(oS).max() < 1e-12
    assert np.abs(Tv - oT).max() / np.abs(oT).max() < 1e-12


def test_sht_round_trip_lmax383():
    """Config-2 shapes (l_max=383, n_theta=576, n_phi=1152), reduced batch: round trip."""
    torch = _torch()
    from magic_amd.params import Params
    from magic_amd.device import Resolution
    from magic_amd.sht import SHT
    p = Params({"l_max": 383, "n_r_max": 9})
    assert (p.n_theta, p.n_phi, p.lm_max) == (576, 1152, 73920)
    res = Resolution(p)
    sht = SHT(res)
    S = _rand_spec(p.lm_max, 6, res.maps, 5)
    dS = torch.from_numpy(S).cuda()
    back = sht.scal_to_SH(sht.scal_to_spat(dS)).cpu().numpy()
    assert np.abs(back - S).max() / np.abs(S).max() < 1e-11
    T = _rand_spec(p.lm_max, 6, res.maps, 6)
    T[0] = 0.0
    S[0] = 0.0
    Z = np.zeros_like(S)
    _, vt, vp = sht.torpol_to_spat(torch.from_numpy(Z).cuda(), torch.from_numpy(S).cuda(),
                                   torch.from_numpy(T).cuda())
    S2, T2 = (x.cpu().numpy() for x in sht.spat_to_sphertor(vt, vp))
    assert np.abs(S2 - S).max() / np.abs(S).max() < 1e-11
    assert np.abs(T2 - T).max() / np.abs(T).max() < 1e-11


@pytest.mark.parametrize("l_max", [63, 383])
def test_sht_mixed_precision_against_fp64(l_max):
    """FP32-MFMA Legendre transforms ("mixed") against the FP64 transforms on
    (l+1)^-2 spectra (the config-2 input law): each of the four transforms and
    the round trip stay within 1e-5 of the largest value (FP32 tables, inputs
    rounded to FP32, FP32 accumulation over at most l_max/2 degrees or n_theta/2
    nodes).  The FP64 transforms are pinned against the oracle above."""
    torch = _torch()
    from magic_amd.params import Params
    from magic_amd.device import Resolution
    from magic_amd.sht import SHT
    p = Params({"l_max": l_max, "n_r_max": 9})
    res = Resolution(p)
    s64, s32 = SHT(res), SHT(res, precision="mixed")
    rel = lambda a, b: float((a - b).abs().max() / b.abs().max())
    S = torch.from_numpy(_rand_spec(p.lm_max, 5, res.maps, 7)).cuda()
    T = torch.from_numpy(_rand_spec(p.lm_max, 5, res.maps, 8)).cuda()
    Q = torch.from_numpy(_rand_spec(p.lm_max, 5, res.maps, 9)).cuda()
    g64 = s64.scal_to_spat(S)
    assert rel(s32.scal_to_spat(S), g64) < 1e-5
    assert rel(s32.scal_to_SH(g64), s64.scal_to_SH(g64)) < 1e-5
    v64 = s64.torpol_to_spat(Q, S, T)
    for a, b in zip(s32.torpol_to_spat(Q, S, T), v64):
        assert rel(a, b) < 1e-5
    for a, b in zip(s32.spat_to_sphertor(v64[1], v64[2]), s64.spat_to_sphertor(v64[1], v64[2])):
        assert rel(a, b) < 1e-5
    assert rel(s32.scal_to_SH(s32.scal_to_spat(S)), S) < 1e-5


@pytest.mark.parametrize("l_max,precision,tol", [(63, "mixed_bf16", 1e-4), (63, "mixed_bf16x3", 1e-5),
                                                  (383, "mixed_bf16", 1e-4), (383, "mixed_bf16x3", 1e-5)])
def test_sht_bf16_split_against_fp64(l_max, precision, tol):
    """BF16 split-precision Legendre transforms (tables and inputs as two or
    three bf16 planes, the plane products i + j < planes on the BF16 MFMA, FP32
    accumulation) against the FP64 transforms on (l+1)^-2 spectra: every
    transform, both directions, scalar and vector, within `tol` of the largest
    value.  Two planes keep about 16 significant bits per operand (dropped
    products below 2^-16 relative), three about 24 (FP32 inputs)."""
    torch = _torch()
    from magic_amd.params import Params
    from magic_amd.device import Resolution
    from magic_amd.sht import SHT
    p = Params({"l_max": l_max, "n_r_max": 9})
    res = Resolution(p)
    s64, sbf = SHT(res), SHT(res, precision=precision)
    rel = lambda a, b: float((a - b).abs().max() / b.abs().max())
    S = torch.from_numpy(_rand_spec(p.lm_max, 5, res.maps, 7)).cuda()
    T = torch.from_numpy(_rand_spec(p.lm_max, 5, res.maps, 8)).cuda()
    Q = torch.from_numpy(_rand_spec(p.lm_max, 5, res.maps, 9)).cuda()
    g64 = s64.scal_to_spat(S)
    errs = [rel(sbf.scal_to_spat(S), g64), rel(sbf.scal_to_SH(g64), s64.scal_to_SH(g64))]
    v64 = s64.torpol_to_spat(Q, S, T)
    errs += [rel(a, b) for a, b in zip(sbf.torpol_to_spat(Q, S, T), v64)]
    errs += [rel(a, b) for a, b in zip(sbf.spat_to_sphertor(v64[1], v64[2]),
                                       s64.spat_to_sphertor(v64[1], v64[2]))]
    errs.append(rel(sbf.scal_to_SH(sbf.scal_to_spat(S)), S))
    print(f"bf16 split {precision} l_max={l_max}: max errors {['%.1e' % e for e in errs]}")
    assert max(errs) < tol, errs
    # the split is doing its work: an error far below one bf16 rounding (2^-9)
    assert max(errs) < 2.0 ** -12


@pytest.mark.parametrize("n", [5, 97, 112, 113, 194, 208, 230])
def test_gj_inverse_matches_numpy(n):
    """Batched equilibrated Gauss-Jordan inverse (register-resident for n <= 208,
    global-memory kernel above) against numpy.linalg.inv, for every
    equilibration mode, on badly scaled random matrices; the batch is done in
    place (A and out alias), as the operator rebuild calls it."""
    torch = _torch()
    from magic_amd import _lib
    from magic_amd.device import ptr, stream_handle
    rng = np.random.default_rng(n)
    batch = 3
    A = rng.standard_normal((batch, n, n)) + n * np.eye(n) * (rng.random((batch, 1, 1)) > 0.5)
    A *= np.exp(rng.uniform(-6, 6, (batch, n, 1)))  # row scales
    A *= np.exp(rng.uniform(-3, 3, (batch, 1, n)))  # column scales
    ref = np.linalg.inv(A)
    for precond in (0, 1, 2):
        d = torch.from_numpy(A.copy()).cuda()
        work = torch.empty(batch * 2 * n * n, dtype=torch.float64, device="cuda")
        scal = torch.empty(batch * 2 * n, dtype=torch.float64, device="cuda")
        info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
        _lib.call("mh_gj_inverse", ptr(d), ptr(d), ptr(work), ptr(scal), n, batch, precond,
                  ptr(info), None, stream_handle())
        got = d.cpu().numpy()
        assert np.all(info.cpu().numpy() == 0)
        for b in range(batch):
            cond = np.linalg.cond(A[b])
            res = np.abs(got[b] @ A[b] - np.eye(n)).max()
            assert res < 1e-14 * cond * n, (n, precond, b, res, cond)
            err = np.abs(got[b] - ref[b]).max() / np.abs(ref[b]).max()
            assert err < 1e-15 * cond * n, (n, precond, b, err, cond)


def test_gj_inverse_reports_singular_pivot():
    """A matrix with a zero column reports the first zero pivot (1-based)."""
    torch = _torch()
    from magic_amd import _lib
    from magic_amd.device import ptr, stream_handle
    for n in (9, 150):
        A = np.random.default_rng(3).standard_normal((1, n, n))
        A[0, :, 0] = 0.0
        d = torch.from_numpy(A).cuda()
        work = torch.empty(2 * n * n, dtype=torch.float64, device="cuda")
        scal = torch.empty(2 * n, dtype=torch.float64, device="cuda")
        info = torch.zeros(1, dtype=torch.int32, device="cuda")
        _lib.call("mh_gj_inverse", ptr(d), ptr(d), ptr(work), ptr(scal), n, 1, 1, ptr(info),
                  None, stream_handle())
        assert int(info.cpu()[0]) == 1


def test_sht_round_trip_config2_full_batch():
    """Config 2 at its full batch: 3 fields x 97 radial levels = 291 columns at
    l_max=383 (n_theta=576, n_phi=1152), coefficients ~ N(0,1) (l+1)^-2.  Scalar
    and vector round trips in FP64 return the input to 1e-11 of its largest
    coefficient; the mixed-precision (FP32 MFMA) transforms stay within 1e-5
    of the FP64 ones, both for the grid values and after the round trip."""
    torch = _torch()
    from magic_amd.params import Params
    from magic_amd.device import Resolution
    from magic_amd.sht import SHT
    p = Params({"l_max": 383, "n_r_max": 9})
    res = Resolution(p)
    f64, mix = SHT(res), SHT(res, "mixed")
    B = 291
    S = torch.from_numpy(_rand_spec(p.lm_max, B, res.maps, 21)).cuda()
    S[0] = 0.0
    g64 = f64.scal_to_spat(S)
    gmx = mix.scal_to_spat(S)
    assert tuple(g64.shape) == (B, 576, 1152)
    gs = g64.abs().max()
    assert float((gmx - g64).abs().max() / gs) < 1e-5
    back = f64.scal_to_SH(g64)
    ss = S.abs().max()
    assert float((back - S).abs().max() / ss) < 1e-11
    back_mx = mix.scal_to_SH(gmx)
    assert float((back_mx - back).abs().max() / ss) < 1e-5
    del g64, gmx, back, back_mx
    T = torch.from_numpy(_rand_spec(p.lm_max, B, res.maps, 22)).cuda()
    T[0] = 0.0
    Z = torch.zeros_like(S)
    _, vt, vp = f64.torpol_to_spat(Z, S, T)
    S2, T2 = f64.spat_to_sphertor(vt, vp)
    assert float((S2 - S).abs().max() / ss) < 1e-11
    assert float((T2 - T).abs().max() / T.abs().max()) < 1e-11
    _, vtm, vpm = mix.torpol_to_spat(Z, S, T)
    assert float((vtm - vt).abs().max() / vt.abs().max()) < 1e-5
    S3, T3 = mix.spat_to_sphertor(vtm, vpm)
    assert float((S3 - S2).abs().max() / ss) < 1e-5
    assert float((T3 - T2).abs().max() / T.abs().max()) < 1e-5
