"""mh_gj_inverse_noalloc (the Gauss-Jordan inverse with its work space in a
fixed device array) against mh_gj_inverse (work space passed in), through the
C ABI: bit-identical inverses and pivot flags on two calls in a row, the second
with a smaller input, so the second call runs on the array the first one left."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _systems(rng, batch, n):
    A = rng.standard_normal((batch, n, n)) + n * np.eye(n) * (rng.random((batch, 1, 1)) > 0.5)
    A *= np.exp(rng.uniform(-6, 6, (batch, n, 1)))
    A *= np.exp(rng.uniform(-3, 3, (batch, 1, n)))
    A[-1, :, 3] = 0.0  # one singular system: its first zero pivot is reported
    return A


def test_noalloc_inverse_matches_allocating_entry_point():
    import torch
    from magic_amd import _lib
    from magic_amd.device import ptr, stream_handle
    rng = np.random.default_rng(20261203)
    # generic kernel twice (the second call smaller), then the register kernel
    for n, batch, precond in ((230, 12, 2), (215, 5, 1), (97, 7, 1)):
        assert batch * (2 * n * n + 2 * n) <= _lib.GJ_SCRATCH_DOUBLES
        print(f"gj noalloc case n={n} batch={batch} precond={precond}", flush=True)
        A = torch.from_numpy(_systems(rng, batch, n)).cuda()
        ref = torch.empty_like(A)
        work = torch.empty(batch * 2 * n * n, dtype=torch.float64, device="cuda")
        scal = torch.empty(batch * 2 * n, dtype=torch.float64, device="cuda")
        info_ref = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
        _lib.call("mh_gj_inverse", ptr(A), ptr(ref), ptr(work), ptr(scal), n, batch, precond,
                  ptr(info_ref), None, stream_handle())
        got = torch.full_like(A, float("nan"))
        info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
        _lib.call("mh_gj_inverse_noalloc", ptr(A), ptr(got), n, batch, precond, ptr(info), None,
                  stream_handle())
        torch.cuda.synchronize()
        assert torch.equal(info, info_ref)
        assert int(info[-1]) > 0 and bool((info[:-1] == 0).all())
        assert torch.equal(got, ref)
        # and the inverses are inverses (the nonsingular ones)
        a, g = A[:-1].cpu().numpy(), got[:-1].cpu().numpy()
        for b in range(batch - 1):
            res = np.abs(g[b] @ a[b] - np.eye(n)).max()
            assert res < 1e-14 * np.linalg.cond(a[b]) * n, (n, b, res)


def test_noalloc_inverse_refuses_sizes_beyond_its_work_space():
    from magic_amd import _lib
    from magic_amd.device import stream_handle
    n = 300
    batch = _lib.GJ_SCRATCH_DOUBLES // (2 * n * n + 2 * n) + 1
    with pytest.raises(_lib.MagicHipError, match="work space"):
        _lib.call("mh_gj_inverse_noalloc", None, None, n, batch, 1, None, None, stream_handle())
