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
  const long lma = va ? (J.rows ? J.rows[ra] : ra) : 0;
  const double* xrow = J.X + lma * 2L * J.ldx + (c16 & 1);
  double4_t acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = (double4_t){0.0, 0.0, 0.0, 0.0};
  const double* mt = J.MT + col0 + c16;
  // operands of k-step k0 (4 k values); the next step's are loaded before the
  // current step's MFMAs are issued
  auto load = [&](int k0, double& av, double (&bv)[4]) {
    const int k = k0 + kk;
    av = (va && k < J.K) ? xrow[2 * k] : 0.0;
    const double* mk = mt + (long)k * J.Rpad;  // MT rows are padded to a multiple of 4
#pragma unroll
    for (int q = 0; q < 4; ++q) bv[q] = mk[16 * q];
  };
  double av, bv[4];
  load(0, av, bv);
  for (int k0 = 0; k0 < J.K; k0 += 4) {
    double an = av, bn[4] = {bv[0], bv[1], bv[2], bv[3]};
    if (k0 + 4 < J.K) load(k0 + 4, an, bn);
#pragma unroll
    for (int q = 0; q < 4; ++q) acc[q] = mfma_f64(av, bv[q], acc[q]);
    av = an;
#pragma unroll
    for (int q = 0; q < 4; ++q) bv[q] = bn[q];
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int row = mfma_row(lane, i);  // 16 rows = 8 lm rows x (re, im)
    const int ro = rbase + (row >> 1);
    if (ro >= J.nrows) continue;
    const long lmo = J.rows ? J.rows[ro] : ro;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = col0 + 16 * q + c16;
      if (r < J.R) {
        const int seg = r / J.rseg, rc = r - seg * J.rseg;
        J.Y[seg][lmo * 2L * J.ldy + 2 * rc + (row & 1)] = acc[q][i];
      }
    }
  }
}

extern "C" int mh_rowop(const MhRowJob* jobs, const int* tiles, int ntiles, void* stream) {
  if (ntiles < 0) return MH_EINVAL;
  if (ntiles == 0) return MH_OK;
  hipLaunchKernelGGL(k_rowop, dim3(ntiles), dim3(256), 0, mh_stream(stream), jobs, tiles);
  return mh_launch_status();
}

