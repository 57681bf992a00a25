// Row operators on spectral fields: Y[lm][r] = sum_k M_g[r][k] X[lm][k] for all
// rows lm of a group g (one degree l for the implicit solves, or every lm for
// a radial derivative), complex X/Y, real M.  One launch walks a tile list
// over many (group, operator) jobs, so all per-degree systems of uneven size
// (l+1 orders per degree) and all their right-hand sides go in one launch.
//
// M is stored transposed and padded: MT[k][r], k < roundup(K,4), r < Rpad
// (multiple of 64), zeros in the padding.  Output column r goes to
// Y[r / rseg] at column r % rseg, which lets one fused operator emit a field
// and its radial derivatives into separate arrays.
#include "common.h"

typedef MhRowJob RowJob;

// The job descriptors live in global memory, so the compiler cannot tell that
// their pointers are global: without this cast every access is a flat access,
// which counts against both wait counters and forces a full drain per step.
#define MH_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const MH_GLOBAL T* gconst(const T* p) {
  return (const MH_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ MH_GLOBAL T* gmut(T* p) {
  return (MH_GLOBAL T*)p;
}

// tiles: (job, row0, col0) triples; WG = 4 waves x (8 rows, 64 columns).
__global__ __launch_bounds__(256, 4) void k_rowop(const RowJob* __restrict__ jobs,
                                               const int* __restrict__ tiles) {
  const int* tl = tiles + 3 * blockIdx.x;
  const RowJob& J = jobs[tl[0]];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int rbase = tl[1] + 8 * w;
  const int col0 = tl[2];
  const int kk = lane >> 4, c16 = lane & 15;
  // A operand: row = c16 -> (row index rbase + c16/2, re/im c16&1)
  const int ra = rbase + (c16 >> 1);
  const bool va = ra < J.nrows;
  const int lma = va ? (J.rows ? gconst(J.rows)[ra] : ra) : 0;
  // uniform bases (scalar registers) and 32-bit per-lane element offsets
  // (a field or RHS array holds fewer than 2^31 doubles)
  const MH_GLOBAL double* xb = gconst(J.X);
  const int xo = lma * 2 * J.ldx + (c16 & 1);
  double4_t acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = (double4_t){0.0, 0.0, 0.0, 0.0};
  const MH_GLOBAL double* mtb = gconst(J.MT) + col0;
  const int mo = kk * J.Rpad + c16;
  // operands of k-step t (k = 4t + kk) stream through a ring of D steps: the
  // loads are unconditional (k clamped to the last row; MT rows are padded to
  // a multiple of 4) and a slot is refilled right after its MFMAs, so each
  // MFMA waits only for loads issued D steps earlier; rows past K or past the
  // job's rows are zeroed where the value is used
  constexpr int D = 4;
  const int nst = (J.K + 3) / 4;
  double xa[D], rb[D][4];
  auto load = [&](int t, double& av, double (&bv)[4]) __attribute__((always_inline)) {
    av = xb[xo + 2 * min(4 * t + kk, J.K - 1)];
    const int m = mo + 4 * t * J.Rpad;  // t is uniform: one scalar multiply per step
#pragma unroll
    for (int q = 0; q < 4; ++q) bv[q] = mtb[m + 16 * q];
  };
#pragma unroll
  for (int d = 0; d < D; ++d) load(min(d, nst - 1), xa[d], rb[d]);
  auto step = [&](int t, int d) __attribute__((always_inline)) {
    const double a = (va && 4 * t + kk < J.K) ? xa[d] : 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = mfma_f64(a, rb[d][q], acc[q]);
    load(min(t + D, nst - 1), xa[d], rb[d]);
  };
  int t0 = 0;
  for (; t0 + D <= nst; t0 += D) {  // full groups: straight-line steps
#pragma unroll
    for (int d = 0; d < D; ++d) step(t0 + d, d);
  }
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (t0 + d < nst) step(t0 + d, d);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int row = mfma_row(lane, i);  // 16 rows = 8 lm rows x (re, im)
    const int ro = rbase + (row >> 1);
    if (ro >= J.nrows) continue;
    const long lmo = J.rows ? gconst(J.rows)[ro] : ro;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = col0 + 16 * q + c16;
      if (r < J.R) {
        const int seg = r / J.rseg, rc = r - seg * J.rseg;
        gmut(J.Y[seg])[lmo * 2L * J.ldy + 2 * rc + (row & 1)] = acc[q][i];
      }
    }
  }
}

// Wide tiles: (job, row0, col0) with col0 a multiple of 128; WG = 4 waves,
// wave w takes lm rows row0 + 16*(w & 1) .. + 15 (two MFMA row tiles of 8 lm
// rows x re/im) and columns col0 + 64*(w >> 1) .. + 63 (four MFMA column
// tiles), so each A operand feeds four MFMAs and each B operand two: 6 loads
// per 8 MFMAs instead of 5 per 4.  Same arithmetic order per output as
// k_rowop (k ascending in steps of 4), hence the same results bit for bit.
__global__ __launch_bounds__(256, 4) void k_rowop_wide(const RowJob* __restrict__ jobs,
                                                    const int* __restrict__ tiles) {
  const int* tl = tiles + 3 * blockIdx.x;
  const RowJob& J = jobs[tl[0]];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int rbase = tl[1] + 16 * (w & 1);
  const int col0 = tl[2] + 64 * (w >> 1);
  // no barriers in this kernel: a wave with no rows or past the padded
  // operator width (a multiple of 64) has nothing to do
  if (rbase >= J.nrows || col0 >= J.Rpad) return;
  const int kk = lane >> 4, c16 = lane & 15;
  const MH_GLOBAL double* xb = gconst(J.X);
  int xo[2];
  bool va[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const int ra = rbase + 8 * i + (c16 >> 1);
    va[i] = ra < J.nrows;
    const int lma = va[i] ? (J.rows ? gconst(J.rows)[ra] : ra) : 0;
    xo[i] = lma * 2 * J.ldx + (c16 & 1);
  }
  double4_t acc[2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[i][q] = (double4_t){0.0, 0.0, 0.0, 0.0};
  const MH_GLOBAL double* mtb = gconst(J.MT) + col0;
  const int mo = kk * J.Rpad + c16;
  constexpr int D = 4;
  const int nst = (J.K + 3) / 4;
  double xa[D][2], rb[D][4];
  auto load = [&](int t, double (&av)[2], double (&bv)[4]) __attribute__((always_inline)) {
    const int k = 2 * min(4 * t + kk, J.K - 1);
    av[0] = xb[xo[0] + k];
    av[1] = xb[xo[1] + k];
    const int m = mo + 4 * t * J.Rpad;
#pragma unroll
    for (int q = 0; q < 4; ++q) bv[q] = mtb[m + 16 * q];
  };
#pragma unroll
  for (int d = 0; d < D; ++d) load(min(d, nst - 1), xa[d], rb[d]);
  auto step = [&](int t, int d) __attribute__((always_inline)) {
    const bool kin = 4 * t + kk < J.K;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const double a = (va[i] && kin) ? xa[d][i] : 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[i][q] = mfma_f64(a, rb[d][q], acc[i][q]);
    }
    load(min(t + D, nst - 1), xa[d], rb[d]);
  };
  int t0 = 0;
  for (; t0 + D <= nst; t0 += D) {
#pragma unroll
    for (int d = 0; d < D; ++d) step(t0 + d, d);
  }
#pragma unroll
  for (int d = 0; d < D; ++d)
    if (t0 + d < nst) step(t0 + d, d);
#pragma unroll
  for (int i = 0; i < 2; ++i) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = mfma_row(lane, v);  // 16 rows = 8 lm rows x (re, im)
      const int ro = rbase + 8 * i + (row >> 1);
      if (ro >= J.nrows) continue;
      const long lmo = J.rows ? gconst(J.rows)[ro] : ro;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int r = col0 + 16 * q + c16;
        if (r < J.R) {
          const int seg = r / J.rseg, rc = r - seg * J.rseg;
          gmut(J.Y[seg])[lmo * 2L * J.ldy + 2 * rc + (row & 1)] = acc[i][q][v];
        }
      }
    }
  }
}

extern "C" int mh_rowop(const MhRowJob* jobs, const int* tiles, int ntiles, int wide,
                        void* stream) {
  if (ntiles < 0 || (wide != 0 && wide != 1)) return MH_EINVAL;
  if (ntiles == 0) return MH_OK;
  if (wide)
    hipLaunchKernelGGL(k_rowop_wide, dim3(ntiles), dim3(256), 0, mh_stream(stream), jobs, tiles);
  else
    hipLaunchKernelGGL(k_rowop, dim3(ntiles), dim3(256), 0, mh_stream(stream), jobs, tiles);
  return mh_launch_status();
}

