// Fused grid-space part of the radial loop.  One workgroup takes one theta row
// t and RB consecutive radial levels r0..r0+RB-1 (one "row" each):
//
//   13 inverse FFTs (u, s, B, curl B, curl u in sin(theta)-weighted form)
//   -> nonlinear products and Courant maxima, point by point in registers
//   -> 9 forward FFTs (Advr VSr VxBr Advt Advp VSt VSp VxBt VxBp), scaled by
//      1/n_phi and, for the theta/phi components, 1/sin^2(theta)
//
// Fourier buffers are F[c][m][t][b]: for fixed (c, m, t) the RB levels of a
// workgroup are RB consecutive complex values, so loads and stores move
// RB*16-byte runs.  The grid never leaves the CU.
//
// Real FFT of n_phi = 2*Mh points = complex Stockham FFT of length Mh in LDS
// (even/odd packing).  The last inverse pass and the first forward pass have
// radix RL and run in registers with a fixed thread <-> butterfly map: thread
// (row, q) owns the complex grid points q + i*TP (i < RL, TP = Mh/RL), i.e.
// the real points 2(q+i*TP) and 2(q+i*TP)+1, for every component.  The
// products therefore need no data exchange, and the forward transform starts
// on the values the thread already holds.
//
// Register budget: at most 7 components of RL complex values are live at
// once (u and B for the induction terms; u or B, the three advection
// accumulators and the component being transformed for the force terms),
// which is why B is transformed twice.
//
// FP32 with a compiled plan (the mixed-precision step) runs k_grid_step3, which
// transforms the three components of each vector together (3*RB LDS rows per
// pass, a third of the workgroup barriers), with the same arithmetic per value.
//
// Level block: by default a workgroup takes the most levels its plan fits
// (threads <= 768, LDS); MhGridArgs.rb asks for fewer (1 or 2 levels run
// compiled plans instantiated in gridstep_rb1.hip / gridstep_rb2.hip, other
// counts the run-time plan).  Results do not depend on it; speed does, so the
// tuner picks it per resolution.
#include "gridstep_kernel.h"

__global__ void k_zero_cfl(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

template <typename T>
static void launch_generic(const GridArgs& a, const GPlan& P, dim3 gd, dim3 bd, size_t lds,
                           hipStream_t st) {
  switch (P.RL) {
    case 2: hipLaunchKernelGGL((k_grid_step<0, 2, T, 0>), gd, bd, lds, st, a, P); break;
    case 3: hipLaunchKernelGGL((k_grid_step<0, 3, T, 0>), gd, bd, lds, st, a, P); break;
    case 4: hipLaunchKernelGGL((k_grid_step<0, 4, T, 0>), gd, bd, lds, st, a, P); break;
    default: hipLaunchKernelGGL((k_grid_step<0, 5, T, 0>), gd, bd, lds, st, a, P); break;
  }
}

extern "C" int mh_grid_plan(int Mh, int* out) {
  const GPlan P = make_plan(Mh);
  if (P.RL == 0 || Mh % 2 != 0) return MH_EINVAL;
  out[0] = P.RL;
  out[1] = P.TP;
  out[2] = P.RB;
  return MH_OK;
}

// Whether Mh has a compiled plan (the FP32 launch of a compiled plan
// transforms three components together and needs three times the LDS rows).
static bool grid_compiled(int Mh) {
  switch (Mh) {
#define MH_CCASE(V) case V:
    MH_GRID_COMPILED(MH_CCASE)
#undef MH_CCASE
    return true;
    default: return false;
  }
}

extern "C" int mh_grid_step(const MhGridArgs* args, void* stream) {
  const GridArgs& a = *args;
  if (a.N < 3 || a.T < 1 || a.Mh < 2 || a.M < 1 || a.M > a.Mh) return MH_EINVAL;
  GPlan P = make_plan(a.Mh);
  if (P.RL == 0) return MH_EINVAL;
  // level block: the plan's own (the most levels that fit), or a.rb fewer
  const int own_rb = P.RB;
  if (a.rb < 0 || a.rb > own_rb) return MH_EINVAL;
  const int rb = a.rb > 0 ? a.rb : own_rb;
  P.RB = rb;
  P.threads = (rb * P.TP + 63) / 64 * 64;
  if (rb == own_rb) P.threads = make_plan(a.Mh).threads;
  // every thread covers its share of the (Mh/2 + 1) * RB conjugate-pair items
  if (a.Mh % 2 != 0 || (long)P.threads * ((P.RL + 2) / 2) < (long)(a.Mh / 2 + 1) * P.RB)
    return MH_EINVAL;
  if (a.fprec != 0 && a.fprec != 1) return MH_EINVAL;
  // the kernel forms buffer offsets in 32 bits with 24-bit multiplies: one
  // order's stride (T*4N*2 elements in Fv, the widest) below 2^24 and each
  // component (M orders) below 2^31 elements
  if ((long)a.T * 4 * a.N * 2 >= (1L << 24) || (long)a.M * a.T * 4 * a.N * 2 >= (1L << 31))
    return MH_EINVAL;
  const size_t esz = a.fprec ? sizeof(float2) : sizeof(double2);
  const bool batched =
      a.fprec && !a.generic && (rb == own_rb || rb == 1 || rb == 2) && grid_compiled(a.Mh);
  const size_t lds = (2ull * (batched ? 3 : 1) * P.RB * (a.Mh + 2) + 2ull * a.Mh) * esz;
  if (lds > (size_t)kLdsBytes) return MH_EINVAL;
  const int nRB = (a.N + P.RB - 1) / P.RB;
  const long nwg = (long)a.T * nRB;
  const long grid = ((nwg + 7) / 8) * 8;
  if (grid > 0x7fffffffL) return MH_EINVAL;
  hipLaunchKernelGGL(k_zero_cfl, dim3(mh_cdiv(2 * a.N, 256)), dim3(256), 0, mh_stream(stream),
                     a.cfl, 2 * a.N);
  const dim3 gd((unsigned)grid), bd(P.threads);
  hipStream_t st = mh_stream(stream);
  bool done = false;
  if (!a.generic) {
    if (rb == own_rb)
      done = a.fprec ? launch_compiled<0, float>(a, P, gd, bd, lds, st)
                     : launch_compiled<0, double>(a, P, gd, bd, lds, st);
    else if (rb == 1)
      done = grid_launch_rb1(a, P, gd, bd, lds, st);
    else if (rb == 2)
      done = grid_launch_rb2(a, P, gd, bd, lds, st);
  }
  if (!done) {
    if (a.fprec) launch_generic<float>(a, P, gd, bd, lds, st);
    else launch_generic<double>(a, P, gd, bd, lds, st);
  }
  return mh_launch_status();
}
