"""Dense and banded LU restated from ``algebra.py`` (oracle).

* ``prepare_mat`` / ``solve_mat``: Gaussian elimination with row pivoting by
  the first maximum of |a| in the column, 1/diag stored on the diagonal
  (``algebra.py:18-69,71-131,133-195``).
* ``prepare_band`` / ``solve_band``: LINPACK dgbfa/dgbsl order with negated
  multipliers and 1-based pivots (``algebra.py:371-457,460-506``).
* ``dense_to_band_storage`` (``algebra.py:642-669``) and ``banded_solve_by_l``
  (``algebra.py:672-723``).

Pinned by ``tests/test_oracle_algebra.py``: solutions against
``numpy.linalg.solve``; band pivot sequence against
``scipy.linalg.lapack.dgbtrf`` (same first-maximum partial-pivot rule); the
reference's own seeded band cases (``tests/test_banded_solvers.py:263-372``:
round trip, dense comparison, multi-RHS, forced pivoting) restated as tests.
"""

from __future__ import annotations

import numpy as np

ZERO_TOLERANCE = 1.0e-15


def prepare_mat(a: np.ndarray):
    """Dense LU, 0-based pivots, 1/diag on the diagonal (``algebra.py:18-69``)."""
    a = np.array(a, np.float64, copy=True)
    n = a.shape[0]
    ip = np.zeros(n, np.int64)
    info = 0
    for k in range(n - 1):
        l = k + int(np.argmax(np.abs(a[k:, k])))
        ip[k] = l
        if abs(a[l, k]) > ZERO_TOLERANCE:
            if l != k:
                a[[k, l], :] = a[[l, k], :]
            a[k + 1:, k] = a[k + 1:, k] / a[k, k]
            a[k + 1:, k + 1:] -= np.outer(a[k + 1:, k], a[k, k + 1:])
        else:
            info = k + 1
    ip[n - 1] = n - 1
    if abs(a[n - 1, n - 1]) <= ZERO_TOLERANCE:
        info = n
    if info > 0:
        return a, ip, info
    d = np.arange(n)
    a[d, d] = 1.0 / a[d, d]
    return a, ip, info


def solve_mat(a: np.ndarray, ip: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Solve with factors from :func:`prepare_mat` (``algebra.py:71-195``).

    Same arithmetic as the reference's pairwise column loop up to summation
    order (forward/back substitution with 1/diag multipliers).
    """
    x = np.array(b, copy=True)
    squeeze = x.ndim == 1
    if squeeze:
        x = x[:, None]
    n = a.shape[0]
    for k in range(n - 1):
        m = ip[k]
        if m != k:
            x[[k, m]] = x[[m, k]]
    for k in range(n - 1):
        x[k + 1:] -= a[k + 1:, k:k + 1] * x[k]
    for k in range(n - 1, -1, -1):
        x[k] = x[k] * a[k, k]
        if k > 0:
            x[:k] -= a[:k, k:k + 1] * x[k]
    return x[:, 0] if squeeze else x


def dense_to_band_storage(A: np.ndarray, kl: int, ku: int) -> np.ndarray:
    """abd[kl+ku+i-j, j] = A[i, j]; first kl rows fill-in (``algebra.py:642-669``)."""
    n = A.shape[0]
    ii, jj = np.nonzero(np.abs(A) > 1e-14)
    bad = (ii < jj - ku) | (ii > jj + kl)
    if np.any(bad):
        i, j = int(ii[bad][0]), int(jj[bad][0])
        raise ValueError(f"dense_to_band_storage: nonzero entry A[{i},{j}]={A[i, j]:.3e} "
                         f"outside band kl={kl}, ku={ku}")
    abd = np.zeros((2 * kl + ku + 1, n), A.dtype)
    for j in range(n):
        for i in range(max(0, j - ku), min(n, j + kl + 1)):
            abd[kl + ku + i - j, j] = A[i, j]
    return abd


def prepare_band(abd: np.ndarray, n: int, kl: int, ku: int):
    """LINPACK dgbfa (``algebra.py:371-457``). Returns (abd, pivot 1-based, info)."""
    abd = np.array(abd, np.float64, copy=True)
    m = kl + ku + 1
    pivot = np.zeros(n, np.int64)
    info = 0
    j0 = ku + 2
    j1 = min(n, m) - 1
    if j1 >= j0:
        for jz in range(j0, j1 + 1):
            i0 = m + 1 - jz
            for i in range(i0, kl + 1):
                abd[i - 1, jz - 1] = 0.0
    jz = j1
    ju = 0
    for k in range(1, n):
        kp1 = k + 1
        jz = jz + 1
        if jz <= n and kl >= 1:
            abd[0:kl, jz - 1] = 0.0
        lm = min(kl, n - k)
        l = int(np.argmax(np.abs(abd[m - 1:m + lm, k - 1]))) + m
        pivot[k - 1] = l + k - m
        if abs(abd[l - 1, k - 1]) > ZERO_TOLERANCE:
            if l != m:
                abd[l - 1, k - 1], abd[m - 1, k - 1] = abd[m - 1, k - 1], abd[l - 1, k - 1]
            t = -1.0 / abd[m - 1, k - 1]
            abd[m:m + lm, k - 1] *= t
            ju = min(max(ju, ku + int(pivot[k - 1])), n)
            mm = m
            for j in range(kp1, ju + 1):
                l -= 1
                mm -= 1
                t = abd[l - 1, j - 1]
                if l != mm:
                    abd[l - 1, j - 1] = abd[mm - 1, j - 1]
                    abd[mm - 1, j - 1] = t
                abd[mm:mm + lm, j - 1] += t * abd[m:m + lm, k - 1]
        else:
            info = k
    pivot[n - 1] = n
    if abs(abd[m - 1, n - 1]) <= ZERO_TOLERANCE:
        info = n
    return abd, pivot, info


def solve_band(abd: np.ndarray, n: int, kl: int, ku: int, pivot: np.ndarray,
               rhs: np.ndarray) -> np.ndarray:
    """LINPACK dgbsl (``algebra.py:460-506``); rhs (n,) or (n, k), real or complex."""
    x = np.array(rhs, copy=True)
    squeeze = x.ndim == 1
    if squeeze:
        x = x[:, None]
    m = ku + kl + 1
    if kl != 0:
        for k in range(1, n):
            lm = min(kl, n - k)
            l = int(pivot[k - 1])
            if l != k:
                x[[k - 1, l - 1]] = x[[l - 1, k - 1]]
            x[k:k + lm] += x[k - 1:k] * abd[m:m + lm, k - 1:k]
    for k in range(n, 0, -1):
        x[k - 1] = x[k - 1] / abd[m - 1, k - 1]
        lm = min(k, m) - 1
        la = m - lm
        lb = k - lm
        if lm > 0:
            t = -x[k - 1:k]
            x[lb - 1:lb - 1 + lm] += t * abd[la - 1:la - 1 + lm, k - 1:k]
    return x[:, 0] if squeeze else x


def banded_solve_by_l(abd_by_l, pivot_by_l, l_index, rhs, n, kl, ku,
                      fac_row_by_l=None, fac_col_by_l=None):
    """Per-l banded solve of all lm rows (``algebra.py:672-723``)."""
    sol = np.zeros_like(rhs)
    for l in range(int(l_index.max()) + 1):
        mask = l_index == l
        if not mask.any():
            continue
        r = rhs[mask]
        if fac_row_by_l is not None:
            r = fac_row_by_l[l][None, :] * r
        s = solve_band(abd_by_l[l], n, kl, ku, pivot_by_l[l], r.T).T
        if fac_col_by_l is not None:
            s = fac_col_by_l[l][None, :] * s
        sol[mask] = s
    return sol
