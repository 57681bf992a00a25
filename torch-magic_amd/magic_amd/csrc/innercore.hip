// Conducting inner core: the coupled outer/inner right-hand sides of the B and
// J systems, and the inner-core old states and implicit terms.
//
// The coupled right-hand side of a mode has n_r_tot = N + n_ic entries
// (reference update_b.py:655-668): 0 at the CMB row, the outer IMEX right-hand
// side on rows 1..N-2, 0 on the ICB continuity (N-1) and derivative-match (N)
// rows, then the inner IMEX right-hand side of levels 1..n_ic-1 (inner bulk and
// centre) on rows N+1..N+n_ic-1.  The per-degree coupled operators (inverse
// with the DCT-I and the derivatives of both shells folded in) are applied by
// the row operator.
#include "common.h"

typedef MhIcRhsArgs IcRhsArgs;
typedef MhIcImplArgs IcImplArgs;

#define C2(p) reinterpret_cast<const double2*>(p)
#define W2(p) reinterpret_cast<double2*>(p)

namespace {

// old wimp + impl wl1 + expl0 we0 + expl1 we1 (+ the BPR353 stage terms t4..t7
// weighted by w[0..3]); absent terms are null (set_imex_rhs, time_scheme.py:75-90,240-262)
__device__ __forceinline__ double2 ic_imex(const MhImexSys& s, long i, const double (&wt)[8]) {
  double2 v = cscale(C2(s.old)[i], wt[0]);
  if (s.impl) v = cadd(v, cscale(C2(s.impl)[i], wt[1]));
  if (s.expl0) v = cadd(v, cscale(C2(s.expl0)[i], wt[2]));
  if (s.expl1) v = cadd(v, cscale(C2(s.expl1)[i], wt[3]));
  if (s.t4) v = cadd(v, cscale(C2(s.t4)[i], wt[4]));
  if (s.t5) v = cadd(v, cscale(C2(s.t5)[i], wt[5]));
  if (s.t6) v = cadd(v, cscale(C2(s.t6)[i], wt[6]));
  if (s.t7) v = cadd(v, cscale(C2(s.t7)[i], wt[7]));
  return v;
}

__global__ __launch_bounds__(256) void k_ic_rhs(IcRhsArgs a) {
  double wt[8] = {a.wimp, a.wl1, a.we0, a.we1, 0.0, 0.0, 0.0, 0.0};
  if (a.wdev) {
    for (int q = 0; q < 4; ++q) wt[q] = a.wdev[q];
    for (int q = 0; q < 4; ++q) wt[4 + q] = a.wdev[8 + q];
  }
  const int N = a.N, nt = a.N + a.n_ic;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)a.L * nt) return;
  const int lm = (int)(idx / nt), c = (int)(idx - (long)lm * nt);
  const bool m0 = a.m[lm] == 0;
  double2 vb = make_double2(0.0, 0.0), vj = vb;
  if (c >= 1 && c <= N - 2) {  // outer bulk: the outer right-hand side (already m=0 real)
    vb = C2(a.rb)[(long)lm * N + c];
    vj = C2(a.rj)[(long)lm * N + c];
  } else if (c > N) {          // inner levels 1..n_ic-1
    const long i = (long)lm * a.n_ic + (c - N);
    vb = ic_imex(a.sys[0], i, wt);
    vj = ic_imex(a.sys[1], i, wt);
    if (m0) vb.y = vj.y = 0.0;
  }
  W2(a.out_b)[idx] = vb;
  W2(a.out_j)[idx] = vj;
}

// Inner-core old states dLh or2(r_icb) f and implicit terms
// fac dLh or2(r_icb) (f'' + 2 (l+1) f' / r) on the inner bulk levels and
// fac dLh or2(r_icb) (1 + 2 (l+1)) f'' at the centre (fac = opm / sigma_ratio;
// update_b.py:761-801, 803-873); the mode lm = 0 and the ICB level are left
// as they are.
__global__ __launch_bounds__(256) void k_ic_impl(IcImplArgs a) {
  const int n = a.n_ic;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)a.L * n) return;
  const int lm = (int)(idx / n), i = (int)(idx - (long)lm * n);
  if (lm == 0) return;
  const double dL = a.dLh[lm];
  const double g = dL * a.or2_icb;
  if (a.old_b) {
    W2(a.old_b)[idx] = cscale(C2(a.b)[idx], g);
    W2(a.old_j)[idx] = cscale(C2(a.aj)[idx], g);
  }
  if (i == 0) return;
  const double lp1 = (double)(a.l[lm] + 1);
  const double f = a.fac * g;
  if (i < n - 1) {
    const double c = 2.0 * lp1 * a.O_r_ic[i];
    W2(a.impl_b)[idx] = cscale(cadd(C2(a.ddb)[idx], cscale(C2(a.db)[idx], c)), f);
    W2(a.impl_j)[idx] = cscale(cadd(C2(a.ddj)[idx], cscale(C2(a.dj)[idx], c)), f);
  } else {
    const double c = f * (1.0 + 2.0 * lp1);
    W2(a.impl_b)[idx] = cscale(C2(a.ddb)[idx], c);
    W2(a.impl_j)[idx] = cscale(C2(a.ddj)[idx], c);
  }
}

typedef MhIcRotArgs IcRotArgs;

// Lorentz torque on the inner core and the rigid-rotation induction at the ICB
// (step_time.py:1126-1138,1181-1192): one workgroup; each thread sums its
// grid points, then a tree reduction in LDS.
__global__ __launch_bounds__(1024) void k_ic_torque(IcRotArgs a) {
  __shared__ double red[1024];
  const double om = a.rot[0];
  const double fv = a.or4_icb * a.r_icb * a.r_icb * om;
  const long n = (long)a.T * a.P;
  double acc = 0.0;
  for (long i = threadIdx.x; i < n; i += blockDim.x) {
    const int t = (int)(i / a.P);
    const double br = a.br[i];
    acc += a.gauss[t] * br * a.bp[i];
    a.vt[i] = fv * a.sin2[t] * br;
    a.vp[i] = 0.0;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) a.rot[a.expl_slot] = a.c_lorentz * (a.fac_lt * red[0]);
}

// Inner explicit terms of a rotating core: -omega or2(r_icb) i m l(l+1) f on
// the inner levels 1..n_ic-1, zero at the ICB (update_b.py:875-901).
__global__ __launch_bounds__(256) void k_ic_advect(IcRotArgs a) {
  const int n = a.n_ic;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)a.L * n) return;
  const int lm = (int)(idx / n), i = (int)(idx - (long)lm * n);
  if (i == 0) {
    W2(a.e_b)[idx] = make_double2(0.0, 0.0);
    W2(a.e_j)[idx] = make_double2(0.0, 0.0);
    return;
  }
  const double g = -a.rot[0] * a.or2_icb * (double)a.m[lm] * a.dLh[lm];  // times i
  const double2 b = C2(a.b_ic)[idx], j = C2(a.aj_ic)[idx];
  W2(a.e_b)[idx] = make_double2(-g * b.y, g * b.x);
  W2(a.e_j)[idx] = make_double2(-g * j.y, g * j.x);
}

// omega's IMEX right-hand side (set_imex_rhs_scalar, time_scheme.py:98-108,
// 278-295) as the ICB entry of z(l=1, m=0) (update_z.py:355-371).
__global__ void k_ic_omega_rhs(IcRotArgs a) {
  double w[8] = {1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int q = 0; q < 4; ++q) w[q] = a.wdev[q];
  for (int q = 0; q < 4; ++q) w[4 + q] = a.wdev[8 + q];
  double v = 0.0;
  for (int q = 0; q < 8; ++q)
    if (a.term[q] >= 0) v += w[q] * a.rot[a.term[q]];
  const long row = (long)a.lm10 * a.N;
  W2(a.rhs_z)[row + a.N - 1] = make_double2(v, 0.0);
  W2(a.rhs_z)[row] = make_double2(0.0, 0.0);
}

// omega_ic, the old state and the implicit term of omega's equation from the
// solved z(l=1, m=0) at the ICB (update_z.py:435-501; Boussinesq: visc = 1,
// beta = 0, no gravitational coupling).
__global__ void k_ic_omega_update(IcRotArgs a) {
  const long i = (long)a.lm10 * a.N + a.N - 1;
  const double z10 = C2(a.z)[i].x, dz10 = C2(a.dz)[i].x;
  a.rot[0] = a.c_z10_omega * z10;
  if (a.store_old) a.rot[1] = a.c_dt_z10 * z10;
  a.rot[a.impl_slot] = -(2.0 * a.or1_icb * z10 - dz10);
}

}  // namespace

static int ic_rot_check(const MhIcRotArgs& a) {
  if (!a.rot || a.L < 1 || a.N < 3 || a.lm10 < 0 || a.lm10 >= a.L) return MH_EINVAL;
  if (a.expl_slot < 6 || a.expl_slot > 9 || a.impl_slot < 2 || a.impl_slot > 5) return MH_EINVAL;
  for (int q = 0; q < 8; ++q)
    if (a.term[q] < -1 || a.term[q] > 9) return MH_EINVAL;
  return MH_OK;
}

extern "C" int mh_ic_torque(const MhIcRotArgs* args, void* stream) {
  const IcRotArgs& a = *args;
  if (ic_rot_check(a) != MH_OK || a.T < 1 || a.P < 1 || !a.br || !a.bp || !a.vt || !a.vp ||
      !a.gauss || !a.sin2)
    return MH_EINVAL;
  hipLaunchKernelGGL(k_ic_torque, dim3(1), dim3(1024), 0, mh_stream(stream), a);
  return mh_launch_status();
}

extern "C" int mh_ic_advect(const MhIcRotArgs* args, void* stream) {
  const IcRotArgs& a = *args;
  if (ic_rot_check(a) != MH_OK || a.n_ic < 2 || !a.b_ic || !a.aj_ic || !a.e_b || !a.e_j || !a.m ||
      !a.dLh)
    return MH_EINVAL;
  const long n = (long)a.L * a.n_ic;
  hipLaunchKernelGGL(k_ic_advect, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     mh_stream(stream), a);
  return mh_launch_status();
}

extern "C" int mh_ic_omega_rhs(const MhIcRotArgs* args, void* stream) {
  const IcRotArgs& a = *args;
  if (ic_rot_check(a) != MH_OK || !a.wdev || !a.rhs_z) return MH_EINVAL;
  hipLaunchKernelGGL(k_ic_omega_rhs, dim3(1), dim3(1), 0, mh_stream(stream), a);
  return mh_launch_status();
}

extern "C" int mh_ic_omega_update(const MhIcRotArgs* args, void* stream) {
  const IcRotArgs& a = *args;
  if (ic_rot_check(a) != MH_OK || !a.z || !a.dz) return MH_EINVAL;
  hipLaunchKernelGGL(k_ic_omega_update, dim3(1), dim3(1), 0, mh_stream(stream), a);
  return mh_launch_status();
}

extern "C" int mh_ic_rhs(const MhIcRhsArgs* args, void* stream) {
  const IcRhsArgs& a = *args;
  if (a.L < 1 || a.N < 3 || a.n_ic < 2 || !a.rb || !a.rj || !a.out_b || !a.out_j || !a.m)
    return MH_EINVAL;
  for (int k = 0; k < 2; ++k) {
    if (!a.sys[k].old) return MH_EINVAL;
    if ((a.sys[k].t4 || a.sys[k].t5 || a.sys[k].t6 || a.sys[k].t7) && !a.wdev) return MH_EINVAL;
  }
  const long n = (long)a.L * (a.N + a.n_ic);
  hipLaunchKernelGGL(k_ic_rhs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     mh_stream(stream), a);
  return mh_launch_status();
}

extern "C" int mh_ic_impl(const MhIcImplArgs* args, void* stream) {
  const IcImplArgs& a = *args;
  if (a.L < 1 || a.n_ic < 2 || !a.b || !a.db || !a.ddb || !a.aj || !a.dj || !a.ddj ||
      !a.impl_b || !a.impl_j || !a.dLh || !a.l || !a.O_r_ic || (!a.old_b != !a.old_j))
    return MH_EINVAL;
  const long n = (long)a.L * a.n_ic;
  hipLaunchKernelGGL(k_ic_impl, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     mh_stream(stream), a);
  return mh_launch_status();
}
