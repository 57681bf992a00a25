"""Per-degree and banded solver entry points with the reference's signatures.

* ``chunked_solve_complex(inv_by_l, l_index, rhs)`` — reference
  ``algebra.py:271-310``: every row lm of ``rhs`` is multiplied by the inverse
  of its degree, all degrees in one row-operator launch (``mh_rowop``).
* ``prepare_band(abd, n, kl, ku)`` / ``solve_band_real(abd, n, kl, ku, pivot,
  rhs)`` / ``banded_solve_by_l(...)`` / ``dense_to_band_storage(A, kl, ku)`` —
  reference ``algebra.py:371-506,642-723``: LINPACK dgbfa/dgbsl order (first
  maximum pivot, negated multipliers, 1-based pivots, ``info`` = first zero
  pivot), run by ``mh_band_factor`` / ``mh_band_solve`` on the GPU, all right-
  hand sides of all degrees in one launch.

Inputs may live on the CPU; they are moved to the GPU and results are
returned on the GPU.  Shapes and degree indices are checked on the host
before any launch (``ValueError`` for a shape that does not fit,
``IndexError`` for a degree index outside the factor sets, as the reference's
indexing raises), so no kernel reads past a buffer.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .device import C128, F64, I32, ptr, require_gpu, stream_handle


def _dev():
    return require_gpu("cuda")


def _check_index(lidx, L, nl, what):
    if lidx.ndim != 1 or lidx.shape[0] != L:
        raise ValueError(f"{what}: l_index must have one entry per right-hand-side row "
                         f"({L}), got shape {tuple(lidx.shape)}")
    if L and (lidx.min() < 0 or lidx.max() >= nl):
        bad = int(lidx[(lidx < 0) | (lidx >= nl)][0])
        raise IndexError(f"{what}: degree index {bad} out of range for {nl} factor sets")


def chunked_solve_complex(inv_by_l, l_index, rhs_complex, chunk_size=None):
    """(L, n) complex solution of inv_by_l[l_index[lm]] @ rhs[lm] for every row lm."""
    inv = torch.as_tensor(inv_by_l, dtype=F64)
    rhs = torch.as_tensor(rhs_complex)
    lidx = torch.as_tensor(l_index).cpu().numpy().astype(np.int64)
    if inv.dim() != 3 or inv.shape[1] != inv.shape[2]:
        raise ValueError(f"chunked_solve_complex: inv_by_l must be (nl, n, n), got "
                         f"{tuple(inv.shape)}")
    nl, n, _ = inv.shape
    if rhs.dim() != 2 or rhs.shape[1] != n:
        raise ValueError(f"chunked_solve_complex: rhs must be (L, {n}), got {tuple(rhs.shape)}")
    L = rhs.shape[0]
    _check_index(lidx, L, nl, "chunked_solve_complex")
    dev = _dev()
    inv = inv.to(dev)
    rhs = rhs.to(dev).to(C128).contiguous()
    K4 = -(-n // 4) * 4
    Rp = -(-n // 64) * 64
    MT = torch.zeros(nl, K4, Rp, dtype=F64, device=dev)
    MT[:, :n, :n] = inv.transpose(1, 2)
    order = np.argsort(lidx, kind="stable").astype(np.int32)
    counts = np.bincount(lidx, minlength=nl)
    offs = np.concatenate([[0], np.cumsum(counts)])
    rows = torch.as_tensor(order).to(dev)
    out = torch.zeros_like(rhs)
    jobs, tiles = [], []
    for l in range(nl):
        if counts[l] == 0:
            continue
        j = _lib.RowJob()
        j.X, j.rows, j.MT = ptr(rhs), ptr(rows) + 4 * int(offs[l]), ptr(MT) + 8 * l * K4 * Rp
        j.Y[0] = ptr(out)
        j.ldx, j.ldy, j.K, j.R, j.Rpad, j.rseg, j.nrows = n, n, n, n, Rp, n, int(counts[l])
        for r0 in range(0, j.nrows, 32):
            for c0 in range(0, n, 64):
                tiles.append((len(jobs), r0, c0))
        jobs.append(j)
    arr = (_lib.RowJob * len(jobs))(*jobs)
    dj = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8).to(dev)
    dt = torch.as_tensor(np.asarray(tiles, np.int32).reshape(-1)).to(dev)
    _lib.call("mh_rowop", ptr(dj), ptr(dt), len(tiles), 0, stream_handle())
    torch.cuda.current_stream().synchronize()
    return out


def dense_to_band_storage(A_dense, kl: int, ku: int):
    """abd[kl+ku+i-j, j] = A[i, j]; raises ValueError outside the band (``algebra.py:642-669``)."""
    A = torch.as_tensor(A_dense, dtype=F64).cpu().numpy()
    n = A.shape[0]
    ii, jj = np.nonzero(np.abs(A) > 1e-14)
    bad = (ii < jj - ku) | (ii > jj + kl)
    if np.any(bad):
        i, j = int(ii[bad][0]), int(jj[bad][0])
        raise ValueError(f"dense_to_band_storage: nonzero entry A[{i},{j}]={A[i, j]:.3e} "
                         f"outside band kl={kl}, ku={ku}")
    abd = np.zeros((2 * kl + ku + 1, n))
    for j in range(n):
        lo, hi = max(0, j - ku), min(n, j + kl + 1)
        abd[kl + ku + np.arange(lo, hi) - j, j] = A[lo:hi, j]
    return torch.as_tensor(abd)


def prepare_band_batched(abd_by_l, n: int, kl: int, ku: int):
    """Factor a batch (B, 2kl+ku+1, n) in place on the GPU; returns (abd, pivot int32, info)."""
    abd = torch.as_tensor(abd_by_l, dtype=F64)
    if abd.dim() == 2:
        abd = abd.unsqueeze(0)
    if abd.dim() != 3 or tuple(abd.shape[1:]) != (2 * kl + ku + 1, n):
        raise ValueError(f"prepare_band: abd must be ({2 * kl + ku + 1}, {n}) per matrix, got "
                         f"{tuple(abd.shape)}")
    dev = _dev()
    abd = abd.to(dev).contiguous().clone()
    nb = abd.shape[0]
    piv = torch.zeros(nb, n, dtype=I32, device=dev)
    info = torch.zeros(nb, dtype=I32, device=dev)
    _lib.call("mh_band_factor", ptr(abd), ptr(piv), ptr(info), n, kl, ku, nb, stream_handle())
    return abd, piv, info


def prepare_band(abd, n: int, kl: int, ku: int):
    """LINPACK dgbfa (``algebra.py:371-457``): (factored abd, 1-based pivots, info)."""
    a, piv, info = prepare_band_batched(abd, n, kl, ku)
    return a[0], piv[0].to(torch.int64), int(info[0].item())


def banded_solve_by_l(abd_by_l, pivot_by_l, l_index, rhs_complex, n, kl, ku,
                      fac_row_by_l=None, fac_col_by_l=None):
    """Every row lm solved with the factors of degree l_index[lm] (``algebra.py:672-723``)."""
    if not (0 <= kl <= 4 and 0 <= ku <= 4):
        raise ValueError(f"banded_solve_by_l: kl, ku <= 4 supported, got {kl}, {ku}")
    abd = torch.as_tensor(abd_by_l, dtype=F64)
    piv = torch.as_tensor(pivot_by_l)
    if abd.dim() != 3 or tuple(abd.shape[1:]) != (2 * kl + ku + 1, n):
        raise ValueError(f"banded_solve_by_l: abd_by_l must be (nl, {2 * kl + ku + 1}, {n}), "
                         f"got {tuple(abd.shape)}")
    nl = abd.shape[0]
    if tuple(piv.shape) != (nl, n):
        raise ValueError(f"banded_solve_by_l: pivot_by_l must be ({nl}, {n}), got "
                         f"{tuple(piv.shape)}")
    lidx_h = torch.as_tensor(l_index).cpu().numpy().astype(np.int64).reshape(-1)
    rhs = torch.as_tensor(rhs_complex)
    if rhs.dim() != 2 or rhs.shape[1] != n:
        raise ValueError(f"banded_solve_by_l: rhs must be (L, {n}), got {tuple(rhs.shape)}")
    _check_index(lidx_h, rhs.shape[0], nl, "banded_solve_by_l")
    for name, fac in (("fac_row_by_l", fac_row_by_l), ("fac_col_by_l", fac_col_by_l)):
        if fac is not None and tuple(torch.as_tensor(fac).shape) != (nl, n):
            raise ValueError(f"banded_solve_by_l: {name} must be ({nl}, {n}), got "
                             f"{tuple(torch.as_tensor(fac).shape)}")
    dev = _dev()
    abd = abd.to(dev).contiguous()
    piv = piv.to(dev).to(I32).contiguous()
    rhs = rhs.to(dev)
    lidx = torch.as_tensor(lidx_h.astype(np.int32)).to(dev).contiguous()
    real = not rhs.is_complex()
    rhs = rhs.to(C128).contiguous()
    fr = None if fac_row_by_l is None else torch.as_tensor(fac_row_by_l, dtype=F64).to(dev).contiguous()
    fc = None if fac_col_by_l is None else torch.as_tensor(fac_col_by_l, dtype=F64).to(dev).contiguous()
    sol = torch.empty_like(rhs)
    rows = torch.argsort(lidx, stable=True).to(I32)   # group the rows by degree
    _lib.call("mh_band_solve", ptr(abd), ptr(piv), ptr(fr), ptr(fc), ptr(lidx), ptr(rows),
              ptr(rhs), ptr(sol), rhs.shape[0], n, kl, ku, stream_handle())
    return sol.real.contiguous() if real else sol


def solve_band_real(abd, n: int, kl: int, ku: int, pivot, rhs):
    """dgbsl (``algebra.py:460-506``): rhs (n,) or (n, k), real or complex."""
    rhs = torch.as_tensor(rhs)
    if rhs.dim() not in (1, 2) or rhs.shape[0] != n:
        raise ValueError(f"solve_band_real: rhs must be ({n},) or ({n}, k), got {tuple(rhs.shape)}")
    dev = _dev()
    rhs = rhs.to(dev)
    vec = rhs.dim() == 1
    cols = rhs.unsqueeze(1) if vec else rhs
    rows = cols.transpose(0, 1).contiguous()             # (k, n) rows
    lidx = torch.zeros(rows.shape[0], dtype=I32, device=dev)
    a = torch.as_tensor(abd, dtype=F64).to(dev).unsqueeze(0)
    p = torch.as_tensor(pivot).to(dev).unsqueeze(0)
    out = banded_solve_by_l(a, p, lidx, rows, n, kl, ku).transpose(0, 1)
    return out[:, 0] if vec else out
