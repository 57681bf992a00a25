// Pivoted band LU in LINPACK dgbfa/dgbsl order for the finite-difference
// per-degree systems.  Band storage abd[(2kl+ku+1) x n] (row kl+ku is the
// diagonal, the first kl rows hold fill-in), multipliers stored negated,
// pivots 1-based, singular pivot threshold 1e-15.  The factorisation is
// sequential in the column index by nature; one lane factors one matrix.
// The solve runs one lane per right-hand-side row (complex), all rows of all
// degrees in one launch, each row using its degree's factors.
#include "common.h"

#define AB(i, j) abd[(long)(i) * n + (j)]  // 0-based band row i, column j

__global__ void k_band_factor(double* __restrict__ abd_all, int* __restrict__ piv_all,
                              int* __restrict__ info_all, int n, int kl, int ku, int nbatch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbatch) return;
  const int nrow = 2 * kl + ku + 1;
  double* abd = abd_all + (long)b * nrow * n;
  int* pivot = piv_all + (long)b * n;
  const int m = kl + ku + 1;  // 1-based diagonal row
  int info = 0;
  const int j0 = ku + 2, j1 = (n < m ? n : m) - 1;
  for (int jz = j0; jz <= j1; ++jz)
    for (int i = m + 1 - jz; i <= kl; ++i) AB(i - 1, jz - 1) = 0.0;
  int jz = j1, ju = 0;
  for (int k = 1; k <= n - 1; ++k) {
    const int kp1 = k + 1;
    jz = jz + 1;
    if (jz <= n && kl >= 1)
      for (int i = 1; i <= kl; ++i) AB(i - 1, jz - 1) = 0.0;
    const int lm = kl < (n - k) ? kl : (n - k);
    // first maximum of |abd(m..m+lm, k)|
    int l = m;
    double best = fabs(AB(m - 1, k - 1));
    for (int i = 1; i <= lm; ++i) {
      const double v = fabs(AB(m - 1 + i, k - 1));
      if (v > best) { best = v; l = m + i; }
    }
    pivot[k - 1] = l + k - m;
    if (fabs(AB(l - 1, k - 1)) > 1.0e-15) {
      if (l != m) {
        const double t = AB(l - 1, k - 1);
        AB(l - 1, k - 1) = AB(m - 1, k - 1);
        AB(m - 1, k - 1) = t;
      }
      const double t = -1.0 / AB(m - 1, k - 1);
      for (int i = 0; i < lm; ++i) AB(m + i, k - 1) *= t;
      int jmax = ku + pivot[k - 1];
      if (ju > jmax) jmax = ju;
      ju = jmax < n ? jmax : n;
      int mm = m, ll = l;
      for (int j = kp1; j <= ju; ++j) {
        ll -= 1;
        mm -= 1;
        const double t2 = AB(ll - 1, j - 1);
        if (ll != mm) {
          AB(ll - 1, j - 1) = AB(mm - 1, j - 1);
          AB(mm - 1, j - 1) = t2;
        }
        for (int i = 0; i < lm; ++i) AB(mm + i, j - 1) += t2 * AB(m + i, k - 1);
      }
    } else {
      info = k;
    }
  }
  pivot[n - 1] = n;
  if (fabs(AB(m - 1, n - 1)) <= 1.0e-15) info = n;
  info_all[b] = info;
}

extern "C" int mh_band_factor(double* abd, int* pivot, int* info, int n, int kl, int ku,
                              int nbatch, void* stream) {
  if (n < 1 || kl < 0 || ku < 0 || nbatch < 1) return MH_EINVAL;
  hipLaunchKernelGGL(k_band_factor, dim3(mh_cdiv(nbatch, 64)), dim3(64), 0, mh_stream(stream),
                     abd, pivot, info, n, kl, ku, nbatch);
  return mh_launch_status();
}

__global__ void k_band_solve(const double* __restrict__ abd_all, const int* __restrict__ piv_all,
                             const double* __restrict__ fac_row, const double* __restrict__ fac_col,
                             const int* __restrict__ lm2l, const double2* __restrict__ rhs,
                             double2* __restrict__ sol, int L, int n, int kl, int ku) {
  const int lm = blockIdx.x * blockDim.x + threadIdx.x;
  if (lm >= L) return;
  const int l = lm2l[lm];
  const int nrow = 2 * kl + ku + 1;
  const double* abd = abd_all + (long)l * nrow * n;
  const int* pivot = piv_all + (long)l * n;
  const double* fr = fac_row ? fac_row + (long)l * n : nullptr;
  const double* fc = fac_col ? fac_col + (long)l * n : nullptr;
  double2* x = sol + (long)lm * n;
  const double2* b = rhs + (long)lm * n;
  for (int i = 0; i < n; ++i) x[i] = fr ? cscale(b[i], fr[i]) : b[i];
  const int m = ku + kl + 1;
  if (kl != 0) {
    for (int k = 1; k <= n - 1; ++k) {
      const int lmk = kl < (n - k) ? kl : (n - k);
      const int lp = pivot[k - 1];
      if (lp != k) {
        const double2 t = x[k - 1];
        x[k - 1] = x[lp - 1];
        x[lp - 1] = t;
      }
      const double2 t = x[k - 1];
      for (int i = 0; i < lmk; ++i) {
        const double a = AB(m + i, k - 1);
        x[k + i] = make_double2(x[k + i].x + t.x * a, x[k + i].y + t.y * a);
      }
    }
  }
  for (int k = n; k >= 1; --k) {
    const double d = AB(m - 1, k - 1);
    x[k - 1] = make_double2(x[k - 1].x / d, x[k - 1].y / d);
    const int lmk = (k < m ? k : m) - 1;
    const int la = m - lmk, lb = k - lmk;
    const double2 t = make_double2(-x[k - 1].x, -x[k - 1].y);
    for (int i = 0; i < lmk; ++i) {
      const double a = AB(la - 1 + i, k - 1);
      x[lb - 1 + i] = make_double2(x[lb - 1 + i].x + t.x * a, x[lb - 1 + i].y + t.y * a);
    }
  }
  if (fc)
    for (int i = 0; i < n; ++i) x[i] = cscale(x[i], fc[i]);
}

extern "C" int mh_band_solve(const double* abd, const int* pivot, const double* fac_row,
                             const double* fac_col, const int* lm2l, const double* rhs,
                             double* sol, int L, int n, int kl, int ku, void* stream) {
  if (L < 1 || n < 1 || kl < 0 || ku < 0) return MH_EINVAL;
  hipLaunchKernelGGL(k_band_solve, dim3(mh_cdiv(L, 64)), dim3(64), 0, mh_stream(stream), abd,
                     pivot, fac_row, fac_col, lm2l, reinterpret_cast<const double2*>(rhs),
                     reinterpret_cast<double2*>(sol), L, n, kl, ku);
  return mh_launch_status();
}
