// Pivoted band LU in LINPACK dgbfa/dgbsl order for the finite-difference
// per-degree systems.  Band storage abd[(2kl+ku+1) x n] (row kl+ku is the
// diagonal, the first kl rows hold fill-in), multipliers stored negated,
// pivots 1-based, singular pivot threshold 1e-15.  The factorisation is
// sequential in the column index by nature: one wave factors one matrix held
// in LDS (one lane per matrix when the band does not fit in 64 KB).  The solve
// runs all right-hand-side rows of all degrees in one launch, each row using
// its degree's factors, over coalesced LDS tiles.  Band widths up to kl, ku = 4
// (the 9-band double-curl operator) are supported by the solve.
#include "common.h"

#include <type_traits>

#define AB(i, j) abd[(long)(i) * n + (j)]  // 0-based band row i, column j

__global__ void k_band_factor(double* __restrict__ abd_all, int* __restrict__ piv_all,
                              int* __restrict__ info_all, int n, int kl, int ku, int nbatch,
                              const int* when) {
  if (when && *when == 0) return;
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

// The same factorisation with one wave per matrix: the band is held in LDS
// (loaded and stored coalesced), every lane finds the pivot (broadcast reads),
// lanes scale the multipliers of column k and update one later column each.
// Same operations in the same order per element as k_band_factor.
__global__ __launch_bounds__(64) void k_band_factor_lds(double* __restrict__ abd_all,
                                                        int* __restrict__ piv_all,
                                                        int* __restrict__ info_all, int n, int kl,
                                                        int ku, const int* when) {
  if (when && *when == 0) return;
  extern __shared__ double sb[];  // [nrow][n]
  const int b = blockIdx.x, lane = threadIdx.x;
  const int nrow = 2 * kl + ku + 1;
  const int m = kl + ku + 1;  // 1-based diagonal row
  double* g = abd_all + (long)b * nrow * n;
  int* pivot = piv_all + (long)b * n;
#define SB(i, j) sb[(long)(i) * n + (j)]  // 0-based band row i, column j
  for (int e = lane; e < nrow * n; e += 64) sb[e] = g[e];
  __syncthreads();
  const int j0 = ku + 2, j1 = (n < m ? n : m) - 1;
  for (int jz = j0; jz <= j1; ++jz)
    for (int i = m + 1 - jz + lane; i <= kl; i += 64) SB(i - 1, jz - 1) = 0.0;
  int info = 0, jz = j1, ju = 0;
  for (int k = 1; k <= n - 1; ++k) {
    jz = jz + 1;
    __syncthreads();  // the previous step's updates are visible
    if (jz <= n)
      for (int i = lane; i < kl; i += 64) SB(i, jz - 1) = 0.0;
    __syncthreads();
    const int lm = kl < (n - k) ? kl : (n - k);
    // first maximum of |abd(m..m+lm, k)|, found by every lane
    int l = m;
    double best = fabs(SB(m - 1, k - 1));
    for (int i = 1; i <= lm; ++i) {
      const double v = fabs(SB(m - 1 + i, k - 1));
      if (v > best) {
        best = v;
        l = m + i;
      }
    }
    const int pk = l + k - m;
    if (lane == 0) pivot[k - 1] = pk;
    const double pv = SB(l - 1, k - 1), dg = SB(m - 1, k - 1);
    if (fabs(pv) > 1.0e-15) {  // uniform: every lane read the same column
      __syncthreads();         // every lane has read the column before it changes
      if (lane == 0 && l != m) {
        SB(l - 1, k - 1) = dg;
        SB(m - 1, k - 1) = pv;
      }
      __syncthreads();
      const double t = -1.0 / pv;
      for (int i = lane; i < lm; i += 64) SB(m + i, k - 1) *= t;
      int jmax = ku + pk;
      if (ju > jmax) jmax = ju;
      ju = jmax < n ? jmax : n;
      __syncthreads();  // multipliers in place
      for (int j = k + 1 + lane; j <= ju; j += 64) {
        const int ll = l - (j - k), mm = m - (j - k);
        const double t2 = SB(ll - 1, j - 1);
        if (ll != mm) {
          SB(ll - 1, j - 1) = SB(mm - 1, j - 1);
          SB(mm - 1, j - 1) = t2;
        }
        for (int i = 0; i < lm; ++i) SB(mm + i, j - 1) += t2 * SB(m + i, k - 1);
      }
    } else {
      info = k;
    }
  }
  __syncthreads();
  if (lane == 0) {
    pivot[n - 1] = n;
    if (fabs(SB(m - 1, n - 1)) <= 1.0e-15) info = n;
    info_all[b] = info;
  }
  for (int e = lane; e < nrow * n; e += 64) g[e] = sb[e];
#undef SB
}

extern "C" int mh_band_factor_gated(double* abd, int* pivot, int* info, int n, int kl, int ku,
                                    int nbatch, const int* when, void* stream) {
  if (n < 1 || kl < 0 || ku < 0 || nbatch < 1) return MH_EINVAL;
  const size_t lds = (size_t)(2 * kl + ku + 1) * n * sizeof(double);
  if (lds <= 64 * 1024) {
    hipLaunchKernelGGL(k_band_factor_lds, dim3(nbatch), dim3(64), lds, mh_stream(stream), abd,
                       pivot, info, n, kl, ku, when);
  } else {
    hipLaunchKernelGGL(k_band_factor, dim3(mh_cdiv(nbatch, 64)), dim3(64), 0, mh_stream(stream),
                       abd, pivot, info, n, kl, ku, nbatch, when);
  }
  return mh_launch_status();
}

extern "C" int mh_band_factor(double* abd, int* pivot, int* info, int n, int kl, int ku,
                              int nbatch, void* stream) {
  return mh_band_factor_gated(abd, pivot, info, n, kl, ku, nbatch, nullptr, stream);
}

// dgbsl over LDS tiles of 16 columns.  The factors are real, so the real and
// imaginary parts of a complex right-hand side are independent recurrences:
// lane 2r+c runs component c of row r, and a wave owns 32 rows.  Each tile is
// loaded (and stored) cooperatively, sixteen lanes moving one row's 256
// contiguous bytes; each lane then runs its recurrence out of LDS.  Forward
// elimination keeps b(k .. k+kl) in a register window and emits x(k-1) kl
// columns behind the input; the back substitution streams the columns
// downwards with the window x(k-m .. k-1), m = kl+ku+1, emitting x(k-1) m
// columns above the input.  Outputs collect in a ring of LDS tiles that is
// stored as soon as a tile is complete; the forward result is the back
// substitution's input (in `sol`).  `rows` (optional) orders the rows by
// degree so that the lanes of a wave share their factors (broadcast loads).
// In place (rhs == sol) is allowed: every tile is read before any output
// lands on it.
constexpr int kBT = 16;  // columns per tile
constexpr int kBR = 32;  // rows per wave

template <int KL, int KU>
__global__ __launch_bounds__(64) void k_band_solve(const double* __restrict__ abd_all,
                                                   const int* __restrict__ piv_all,
                                                   const double* __restrict__ fac_row,
                                                   const double* __restrict__ fac_col,
                                                   const int* __restrict__ lm2l,
                                                   const int* __restrict__ rows,
                                                   const double2* rhs, double2* sol, int L,
                                                   int n, const int* when) {
  if (when && *when == 0) return;
  constexpr int M = KL + KU + 1;
  constexpr int NR = 2 * KL + KU + 1;
  constexpr int TP = kBT + 1;                    // padded LDS row (conflict-free columns)
  constexpr int NO = (M + kBT - 1) / kBT + 1;    // output ring slots (back substitution)
  constexpr int CR = 64 / kBT;                   // rows per cooperative access
  constexpr int CI = kBR / CR;                   // cooperative accesses per lane
  __shared__ double2 tin[kBR][TP];
  __shared__ double2 tout[NO > 2 ? NO : 2][kBR][TP];
  const int lane = threadIdx.x;
  const int r0 = blockIdx.x * kBR;
  // this lane's row and component, and its degree's factors
  const int row = lane >> 1, comp = lane & 1;
  const int ri = r0 + row;
  const int lm = ri < L ? (rows ? rows[ri] : ri) : 0;
  const int l = lm2l[lm];
  const double* abd = abd_all + (long)l * NR * n;
  const int* pivot = piv_all + (long)l * n;
  const double* fcl = fac_col ? fac_col + (long)l * n : nullptr;
  auto in_at = [&](int c) { return reinterpret_cast<const double*>(&tin[row][c])[comp]; };
  auto out_at = [&](int slot, int pos, double v) {
    reinterpret_cast<double*>(&tout[slot][row][pos])[comp] = v;
  };
  // cooperative map: lane -> sub-row cs (rows cs + CR*it) and column cc of a tile
  static_assert(64 % kBT == 0 && kBR % CR == 0, "cooperative tile map");
  const int cs = lane / kBT, cc = lane % kBT;
  long cbase[CI];
  const double* cfr[CI];
  bool cok[CI];
#pragma unroll
  for (int it = 0; it < CI; ++it) {
    const int rr = r0 + cs + CR * it;
    cok[it] = rr < L;
    const int lmr = cok[it] ? (rows ? rows[rr] : rr) : 0;
    cbase[it] = (long)lmr * n;
    cfr[it] = fac_row ? fac_row + (long)lm2l[lmr] * n : nullptr;
  }
  const int ntile = (n + kBT - 1) / kBT;
  auto coop_load = [&](const double2* src, int tile, bool scale, double2 (&v)[CI]) {
    const int col = tile * kBT + cc;
#pragma unroll
    for (int it = 0; it < CI; ++it) {
      v[it] = make_double2(0.0, 0.0);
      if (cok[it] && col < n) {
        v[it] = src[cbase[it] + col];
        if (scale && cfr[it]) v[it] = cscale(v[it], cfr[it][col]);
      }
    }
  };
  auto coop_put = [&](const double2 (&v)[CI]) {
#pragma unroll
    for (int it = 0; it < CI; ++it) tin[cs + CR * it][cc] = v[it];
  };
  auto coop_store = [&](int slot, int tile) {
    const int col = tile * kBT + cc;
#pragma unroll
    for (int it = 0; it < CI; ++it)
      if (cok[it] && col < n) sol[cbase[it] + col] = tout[slot][cs + CR * it][cc];
  };
  double2 pre[CI];

  // ---- forward elimination: x = L^-1 P b (rows scaled by fac_row)
  if (KL > 0) {
    double W[KL + 1];
#pragma unroll
    for (int j = 0; j <= KL; ++j) W[j] = 0.0;
    // iteration k (1-based): pivot swap, update of b(k+1 .. k+lmk), emit x(k-1)
    // (full: an interior iteration, k <= n - KL, so every update applies)
    auto iter = [&](auto full, int k, int slot, int pos) __attribute__((always_inline)) {
      constexpr bool FULL = decltype(full)::value;
      const int lmk = FULL ? KL : min(KL, n - k);
      const int lp = pivot[k - 1] - k;
      // swap W[0] and W[lp] with selects (an indexed swap would put W in scratch)
      const double w0 = W[0];
      double pv = w0;
#pragma unroll
      for (int j = 1; j <= KL; ++j) {
        const bool s = j == lp;
        pv = s ? W[j] : pv;
        W[j] = s ? w0 : W[j];
      }
      W[0] = pv;
#pragma unroll
      for (int j = 1; j <= KL; ++j)
        if (FULL || j <= lmk) W[j] += pv * abd[(long)(M + j - 1) * n + (k - 1)];
      out_at(slot, pos, W[0]);
#pragma unroll
      for (int j = 0; j < KL; ++j) W[j] = W[j + 1];
      W[KL] = 0.0;
    };
    int flushed = 0;  // output tiles < flushed are stored
    auto ftile = [&](int t) __attribute__((always_inline)) {
      __syncthreads();  // the previous tile's readers are done
      coop_put(pre);
      __syncthreads();
      if (t + 1 < ntile) coop_load(rhs, t + 1, true, pre);
      if (t >= 1 && (t + 1) * kBT <= n) {
        // interior tile: straight-line iterations (their loads can be issued early)
#pragma unroll
        for (int c = 0; c < kBT; ++c) {
          W[KL] = in_at(c);
          const int o = t * kBT + c - KL;
          iter(std::true_type{}, o + 1, (o / kBT) & 1, (c - KL + kBT) % kBT);
        }
      } else {
#pragma unroll
      for (int c = 0; c < kBT; ++c) {
        const int j = t * kBT + c;  // input column entering the window
        if (j < n) {
          W[KL] = in_at(c);
          if (j >= KL) {
            // iteration k = j - KL + 1 emits column j - KL
            const int o = j - KL;
            iter(std::false_type{}, o + 1, (o / kBT) & 1, (c - KL + kBT) % kBT);
          } else {
            // filling: b(0 .. KL-1) move down as they arrive
#pragma unroll
            for (int q = 0; q < KL; ++q) W[q] = W[q + 1];
          }
        }
      }
      }
      // outputs so far cover columns <= hi: store every complete tile (a tile's
      // ring slot is written again only two input tiles later)
      const int hi = min(t * kBT + kBT - 1, n - 1) - KL;
      __syncthreads();
      while (flushed < ntile && flushed * kBT + kBT - 1 <= hi) {
        coop_store(flushed & 1, flushed);
        ++flushed;
      }
    };
    coop_load(rhs, 0, true, pre);
    for (int t = 0; t < ntile; ++t) ftile(t);
    // fewer columns than KL: finish the fill with zeros
    for (int j = n; j < KL; ++j) {
#pragma unroll
      for (int q = 0; q < KL; ++q) W[q] = W[q + 1];
      W[KL] = 0.0;
    }
    // the last iterations (nothing left to enter), then x(n-1) = W[0]
    for (int k = max(1, n - KL + 1); k <= n - 1; ++k) {
      const int o = k - 1;
      iter(std::false_type{}, k, (o / kBT) & 1, o % kBT);
    }
    out_at(((n - 1) / kBT) & 1, (n - 1) % kBT, W[0]);
    __syncthreads();
    while (flushed < ntile) {
      coop_store(flushed & 1, flushed);
      ++flushed;
    }
  }

  // ---- back substitution: x = U^-1 y (columns scaled by fac_col)
  {
    const double2* src = (KL > 0) ? sol : rhs;
    const bool scale_in = (KL == 0);  // without a forward pass the row scaling is applied here
    double V[M];
#pragma unroll
    for (int j = 0; j < M; ++j) V[j] = 0.0;
    // iteration k (1-based, descending): finish x(k-1), update x(k-2 .. k-1-lmk)
    // (full: k > M, so every update applies)
    auto biter = [&](auto full, int k) __attribute__((always_inline)) {
      constexpr bool FULL = decltype(full)::value;
      V[M - 1] /= abd[(long)(M - 1) * n + (k - 1)];
      const int lmk = FULL ? M - 1 : min(k, M) - 1;
      const double t = -V[M - 1];
#pragma unroll
      for (int j = 1; j < M; ++j)
        if (FULL || j <= lmk) V[M - 1 - j] += t * abd[(long)(M - j - 1) * n + (k - 1)];
      const int o = k - 1;
      out_at((o / kBT) % NO, o % kBT, fcl ? V[M - 1] * fcl[o] : V[M - 1]);
#pragma unroll
      for (int j = M - 1; j > 0; --j) V[j] = V[j - 1];
      V[0] = 0.0;
    };
    int flushed = ntile;  // output tiles >= flushed are stored
    auto btile = [&](int u) __attribute__((always_inline)) {
      __syncthreads();
      coop_put(pre);
      __syncthreads();
      if (u > 0) coop_load(src, u - 1, scale_in, pre);
      if ((u + 1) * kBT + M <= n) {
        // interior tile: straight-line iterations (k = col + M + 1 <= n, k > M)
#pragma unroll
        for (int c = kBT - 1; c >= 0; --c) {
          const double v = in_at(c);
          biter(std::true_type{}, u * kBT + c + M + 1);
          V[0] = v;
        }
      } else {
#pragma unroll
      for (int c = kBT - 1; c >= 0; --c) {
        const int col = u * kBT + c;  // input column entering the window from below
        if (col < n) {
          const double v = in_at(c);
          const int k = col + M + 1;
          if (k <= n) {
            biter(std::false_type{}, k);
            V[0] = v;
          } else {
            // initial fill of x(n-1) .. x(n-M), by distance from the top
            const int pos = M - 1 - (n - 1 - col);
#pragma unroll
            for (int j = 0; j < M; ++j)
              if (j == pos) V[j] = v;
          }
        }
      }
      }
      // outputs so far cover columns >= u*kBT + M: store every complete tile
      __syncthreads();
      while (flushed - 1 >= 0 && (flushed - 1) * kBT >= u * kBT + M) {
        --flushed;
        coop_store(flushed % NO, flushed);
      }
    };
    coop_load(src, ntile - 1, scale_in, pre);
    for (int u = ntile - 1; u >= 0; --u) btile(u);
    for (int k = min(M, n); k >= 1; --k) biter(std::false_type{}, k);
    __syncthreads();
    while (flushed - 1 >= 0) {
      --flushed;
      coop_store(flushed % NO, flushed);
    }
  }
}

extern "C" int mh_band_solve_gated(const double* abd, const int* pivot, const double* fac_row,
                                   const double* fac_col, const int* lm2l, const int* rows,
                                   const double* rhs, double* sol, int L, int n, int kl, int ku,
                                   const int* when, void* stream) {
  if (L < 1 || n < 1 || kl < 0 || ku < 0 || kl > 4 || ku > 4) return MH_EINVAL;
  const dim3 g(mh_cdiv(L, kBR)), blk(64);  // one wave of 32 rows (two lanes each) per workgroup
  hipStream_t st = mh_stream(stream);
  const double2* r = reinterpret_cast<const double2*>(rhs);
  double2* o = reinterpret_cast<double2*>(sol);
#define MH_BAND_CASE(A, B)                                                                      \
  if (kl == A && ku == B) {                                                                     \
    hipLaunchKernelGGL((k_band_solve<A, B>), g, blk, 0, st, abd, pivot, fac_row, fac_col, lm2l, \
                       rows, r, o, L, n, when);                                                 \
    return mh_launch_status();                                                                  \
  }
#define MH_BAND_ROW(A) MH_BAND_CASE(A, 0) MH_BAND_CASE(A, 1) MH_BAND_CASE(A, 2) MH_BAND_CASE(A, 3) MH_BAND_CASE(A, 4)
  MH_BAND_ROW(0) MH_BAND_ROW(1) MH_BAND_ROW(2) MH_BAND_ROW(3) MH_BAND_ROW(4)
#undef MH_BAND_ROW
#undef MH_BAND_CASE
  return MH_EINVAL;
}

extern "C" int mh_band_solve(const double* abd, const int* pivot, const double* fac_row,
                             const double* fac_col, const int* lm2l, const int* rows,
                             const double* rhs, double* sol, int L, int n, int kl, int ku,
                             void* stream) {
  return mh_band_solve_gated(abd, pivot, fac_row, fac_col, lm2l, rows, rhs, sol, L, n, kl, ku,
                             nullptr, stream);
}
