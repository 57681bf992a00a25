// Pivoted band LU in LINPACK dgbfa/dgbsl order for the finite-difference
// per-degree systems.  Band storage abd[(2kl+ku+1) x n] (row kl+ku is the
// diagonal, the first kl rows hold fill-in), multipliers stored negated,
// pivots 1-based, singular pivot threshold 1e-15.  The factorisation is
// sequential in the column index by nature; one lane factors one matrix.
// The solve runs one lane per right-hand-side row (complex), all rows of all
// degrees in one launch, each row using its degree's factors.  Band widths
// up to kl, ku = 4 (the 9-band double-curl operator) are supported.
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

// dgbsl for one right-hand-side row per lane, streamed through register
// windows: the forward elimination keeps b(k .. k+kl), the back substitution
// b(k-kl-ku .. k); every element of the row is read and written once per pass.
// `rows` (optional) orders the rows by degree so that the lanes of a wave
// share their factors (broadcast loads).
template <int KL, int KU>
__global__ __launch_bounds__(64) void k_band_solve(const double* __restrict__ abd_all,
                                                   const int* __restrict__ piv_all,
                                                   const double* __restrict__ fac_row,
                                                   const double* __restrict__ fac_col,
                                                   const int* __restrict__ lm2l,
                                                   const int* __restrict__ rows,
                                                   const double2* rhs, double2* sol, int L,
                                                   int n) {
  constexpr int M = KL + KU + 1;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= L) return;
  const int lm = rows ? rows[idx] : idx;
  const int l = lm2l[lm];
  const double* abd = abd_all + (long)l * (2 * KL + KU + 1) * n;
  const int* pivot = piv_all + (long)l * n;
  const double* fr = fac_row ? fac_row + (long)l * n : nullptr;
  const double* fc = fac_col ? fac_col + (long)l * n : nullptr;
  const double2* b = rhs + (long)lm * n;
  double2* x = sol + (long)lm * n;
  auto bin = [&](int i) {  // scaled right-hand side b(i), 0-based, zero past the end
    if (i >= n) return make_double2(0.0, 0.0);
    const double2 v = b[i];
    return fr ? cscale(v, fr[i]) : v;
  };
  // forward elimination, window W[j] = b(k+j), k = 1 .. n (1-based)
  if (KL > 0) {
    double2 W[KL + 1];
#pragma unroll
    for (int j = 0; j <= KL; ++j) W[j] = bin(j);
    for (int k = 1; k <= n - 1; ++k) {
      const int lmk = min(KL, n - k);
      const int lp = pivot[k - 1] - k;  // 0 .. lmk
#pragma unroll
      for (int j = 1; j <= KL; ++j)
        if (j == lp) {
          const double2 t = W[0];
          W[0] = W[j];
          W[j] = t;
        }
      const double2 t = W[0];
#pragma unroll
      for (int j = 1; j <= KL; ++j)
        if (j <= lmk) {
          const double a = AB(M + j - 1, k - 1);
          W[j] = make_double2(W[j].x + t.x * a, W[j].y + t.y * a);
        }
      x[k - 1] = W[0];
#pragma unroll
      for (int j = 0; j < KL; ++j) W[j] = W[j + 1];
      W[KL] = bin(k + KL);
    }
    x[n - 1] = W[0];
  } else {
    for (int i = 0; i < n; ++i) x[i] = bin(i);
  }
  // back substitution, window V[M-1-j] = b(k-j), k = n .. 1
  double2 V[M];
#pragma unroll
  for (int j = 0; j < M; ++j) V[M - 1 - j] = (n - 1 - j >= 0) ? x[n - 1 - j] : make_double2(0.0, 0.0);
  for (int k = n; k >= 1; --k) {
    const double d = AB(M - 1, k - 1);
    V[M - 1] = make_double2(V[M - 1].x / d, V[M - 1].y / d);
    const int lmk = min(k, M) - 1;
    const double2 t = make_double2(-V[M - 1].x, -V[M - 1].y);
#pragma unroll
    for (int j = 1; j < M; ++j)
      if (j <= lmk) {
        const double a = AB(M - j - 1, k - 1);
        V[M - 1 - j] = make_double2(V[M - 1 - j].x + t.x * a, V[M - 1 - j].y + t.y * a);
      }
    x[k - 1] = fc ? cscale(V[M - 1], fc[k - 1]) : V[M - 1];
#pragma unroll
    for (int j = M - 1; j > 0; --j) V[j] = V[j - 1];
    V[0] = (k - M - 1 >= 0) ? x[k - M - 1] : make_double2(0.0, 0.0);
  }
}

extern "C" int mh_band_solve(const double* abd, const int* pivot, const double* fac_row,
                             const double* fac_col, const int* lm2l, const int* rows,
                             const double* rhs, double* sol, int L, int n, int kl, int ku,
                             void* stream) {
  if (L < 1 || n < 1 || kl < 0 || ku < 0 || kl > 4 || ku > 4) return MH_EINVAL;
  const dim3 g(mh_cdiv(L, 64)), blk(64);
  hipStream_t st = mh_stream(stream);
  const double2* r = reinterpret_cast<const double2*>(rhs);
  double2* o = reinterpret_cast<double2*>(sol);
#define MH_BAND_CASE(A, B)                                                                      \
  if (kl == A && ku == B) {                                                                     \
    hipLaunchKernelGGL((k_band_solve<A, B>), g, blk, 0, st, abd, pivot, fac_row, fac_col, lm2l, \
                       rows, r, o, L, n);                                                       \
    return mh_launch_status();                                                                  \
  }
#define MH_BAND_ROW(A) MH_BAND_CASE(A, 0) MH_BAND_CASE(A, 1) MH_BAND_CASE(A, 2) MH_BAND_CASE(A, 3) MH_BAND_CASE(A, 4)
  MH_BAND_ROW(0) MH_BAND_ROW(1) MH_BAND_ROW(2) MH_BAND_ROW(3) MH_BAND_ROW(4)
#undef MH_BAND_ROW
#undef MH_BAND_CASE
  return MH_EINVAL;
}
