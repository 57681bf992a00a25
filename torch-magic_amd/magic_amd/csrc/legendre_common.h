// Helpers shared by the Legendre kernels (legendre.hip: FP64 and FP32 MFMA;
// legendre_bf16.hip: BF16 split-precision MFMA).
#pragma once

#include "common.h"

typedef MhSynIn SynIn;
typedef MhSynArgs SynArgs;
typedef MhAnaArgs AnaArgs;

// Degrees (synthesis) or theta pairs (analysis) staged in LDS per chunk.
constexpr int kChunk = 32;

// i*m*z where lane pairs (2q, 2q+1) hold (re, im) of z
__device__ __forceinline__ double imul_pair(double v, double m, int lane) {
  const double o = __shfl_xor(v, 1);
  return (lane & 1) ? m * o : -m * o;
}

// Offset (in doubles) of the real part of F(c, m, t, b), layout F[c][m][t][b].
__device__ __forceinline__ long fidx(long cs, long ms, int B, int c, int m, int t, int b) {
  return c * cs + m * ms + (long)t * 2 * B + 2L * b;
}

// Per-lane group select without dynamic indexing of the argument struct.
__device__ __forceinline__ SynIn syn_in(const SynArgs& a, int g, int c) {
  SynIn r = a.in[0][c];
  if (g == 1) r = a.in[1][c];
  if (g == 2) r = a.in[2][c];
  if (g == 3) r = a.in[3][c];
  return r;
}


// BF16 split-precision launchers (legendre_bf16.hip), for prec 2 and 3.
int leg_synth_bf16(int vector, const SynArgs& a, int n_m, hipStream_t st);
int leg_analysis_bf16(int vector, const AnaArgs& a, hipStream_t st);
