// Double-curl poloidal equation of the finite-difference step: elementwise
// spectral kernels over (L, N) complex arrays (m-major lm order, lm = 0 is
// l = 0), one thread per (lm, r) point.
//
//  * k_td_dc: explicit poloidal term (get_td.py:91-147 with
//    finish_exp_pol, update_wp.py:371-384); the pressure term is zero.
//  * k_buoy_dc: buoyancy BuoFac dLh/r^2 g s on rows 2..N-3 of the w right-hand
//    side and the l = 0 pressure right-hand side (update_w_doublecurl.py:413-459).
//  * k_impl_dc: old state and implicit terms of w (update_w_doublecurl.py:309-393);
//    the pressure's are zero (no pressure equation).
//  * k_affine: out = A0 + w A1 (operators that are affine in the implicit weight).
#include "common.h"

#define C2(p) reinterpret_cast<const double2*>(p)
#define W2(p) reinterpret_cast<double2*>(p)

typedef MhDcArgs DcArgs;

struct Prof {
  double be, dbe, ddbe, orho, rho, visc, dv, ddv;
};

__device__ __forceinline__ Prof prof_at(const DcArgs& a, int r) {
  Prof q;
  q.be = a.beta ? a.beta[r] : 0.0;
  q.dbe = a.dbeta ? a.dbeta[r] : 0.0;
  q.ddbe = a.ddbeta ? a.ddbeta[r] : 0.0;
  q.orho = a.orho1 ? a.orho1[r] : 1.0;
  q.rho = a.rho0 ? a.rho0[r] : 1.0;
  q.visc = a.visc ? a.visc[r] : 1.0;
  q.dv = a.dLvisc ? a.dLvisc[r] : 0.0;
  q.ddv = a.ddLvisc ? a.ddLvisc[r] : 0.0;
  return q;
}

__global__ __launch_bounds__(256) void k_td_dc(DcArgs a) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int N = a.rc.N;
  if (idx >= (long)a.lc.L * N) return;
  const int lm = (int)(idx / N), r = (int)(idx - (long)lm * N);
  const int l = a.lc.l[lm], m = a.lc.m[lm];
  const bool bdry = (r == 0) || (r == N - 1);
  const double dLh = a.lc.dLh[lm];
  const double or1 = a.rc.or1[r], or2 = a.rc.or2[r], or4 = a.rc.or4[r];
  const double c2 = 2.0 * a.CorFac;
  const bool hasA = l < a.lc.l_max;
  const long iA = hasA ? idx + N : idx;          // (l+1, m)
  const long iS = (l > m) ? idx - N : idx;        // (l-1, m) or self (zero coefficient)
  const double mA = hasA ? 1.0 : 0.0;
  const double2 zA = cscale(C2(a.z)[iA], mA), zS = C2(a.z)[iS];
  double2 e = make_double2(0.0, 0.0);
  if (!bdry) {
    const double2 Advr = C2(a.Advr)[idx];
    if (lm == a.lm_l0) {
      e = cadd(cscale(Advr, or2), cscale(zA, c2 * or1 * a.lc.dTh2A[lm]));
    } else {
      const Prof q = prof_at(a, r);
      const double2 dzA = cscale(C2(a.dz)[iA], mA), dzS = C2(a.dz)[iS];
      const double2 w = C2(a.w)[idx], dw = C2(a.dw)[idx], ddw = C2(a.ddw)[idx];
      double2 ph = csub(cscale(dw, q.be), ddw);
      ph = cadd(ph, cscale(w, (q.be * or1 + or2) * dLh));
      double2 cor = cimul(ph, (double)m);
      cor = cadd(cor, cscale(csub(dzA, cscale(zA, q.be)), a.lc.dTh3A[lm]));
      cor = cadd(cor, cscale(csub(dzS, cscale(zS, q.be)), a.lc.dTh3S[lm]));
      cor = cadd(cor, cscale(csub(cscale(zA, a.lc.dTh4A[lm]), cscale(zS, a.lc.dTh4S[lm])), or1));
      e = cadd(cscale(Advr, dLh * or4 * q.orho), cscale(cor, c2 * or2 * q.orho));
    }
  }
  // + or2 d/dr(-orho1 r^2 dLh Advt)
  e = csub(e, cscale(C2(a.dAdvt_dr)[idx], or2 * dLh));
  W2(a.ew)[idx] = e;
  W2(a.ep)[idx] = make_double2(0.0, 0.0);
}

__global__ __launch_bounds__(256) void k_buoy_dc(DcArgs a) {
  const double fac = a.wdev ? a.wdev[4] : a.fac;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int N = a.rc.N;
  if (idx >= (long)a.lc.L * N) return;
  const int lm = (int)(idx / N), r = (int)(idx - (long)lm * N);
  const long iw = (long)lm * 2 * N + r;
  const double2 s = C2(a.s)[idx];
  if (r >= 2 && r <= N - 3) {
    const double g = fac * a.lc.dLh[lm] * a.rc.or2[r] * a.grav[r];
    double2 v = W2(a.rwp)[iw];
    v.x += g * s.x;
    v.y += (a.lc.m[lm] == 0) ? 0.0 : g * s.y;
    W2(a.rwp)[iw] = v;
  }
  if (lm == a.lm_l0) {
    // l = 0 pressure: 0 at the outer boundary, rho0 BuoFac g Re(s) + Re(dwdt expl) below
    const double rho = a.rho0 ? a.rho0[r] : 1.0;
    const double v = (r == 0) ? 0.0 : rho * a.BuoFac * a.grav[r] * s.x + C2(a.ew0)[idx].x;
    W2(a.rwp)[iw + N] = make_double2(v, 0.0);
  }
}

__global__ __launch_bounds__(256) void k_impl_dc(DcArgs a) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int N = a.rc.N;
  if (idx >= (long)a.lc.L * N) return;
  const int lm = (int)(idx / N), r = (int)(idx - (long)lm * N);
  const bool bdry = (r == 0) || (r == N - 1);
  const double2 zero = make_double2(0.0, 0.0);
  if (a.old_p) W2(a.old_p)[idx] = zero;
  if (a.impl_p) W2(a.impl_p)[idx] = zero;
  if (bdry) {
    if (a.old_w) W2(a.old_w)[idx] = zero;
    W2(a.impl_w)[idx] = zero;
    return;
  }
  const Prof q = prof_at(a, r);
  const double dL = a.lc.dLh[lm];
  const double or1 = a.rc.or1[r], or2 = a.rc.or2[r];
  const double2 w = C2(a.w)[idx], dw = C2(a.dw)[idx], ddw = C2(a.ddw)[idx];
  if (a.old_w) {
    const double2 t = csub(csub(ddw, cscale(dw, q.be)), cscale(w, dL * or2));
    W2(a.old_w)[idx] = cscale(t, -dL * or2 * q.orho);
  }
  const double be = q.be, dbe = q.dbe, ddbe = q.ddbe, dv = q.dv, ddv = q.ddv;
  const double c3 = 2.0 * (dv - be);
  const double c2 = ddv - 2.0 * dbe + dv * dv + be * be - 3.0 * dv * be - 2.0 * or1 * (dv + be) -
                    2.0 * dL * or2;
  const double c1 = -ddbe - dbe * (2.0 * dv - be + 2.0 * or1) - ddv * (be + 2.0 * or1) +
                    be * be * (dv + 2.0 * or1) - be * (dv * dv - 2.0 * or2) -
                    2.0 * dv * or1 * (dv - or1) + 2.0 * (2.0 * or1 + be - dv) * dL * or2;
  const double c0 = dL * or2 *
                    (2.0 * dbe + ddv + dv * dv - (2.0 / 3.0) * be * be + dv * be +
                     2.0 * or1 * (2.0 * dv - be - 3.0 * or1) + dL * or2);
  double2 op = C2(a.ddddw)[idx];
  op = cadd(op, cscale(C2(a.dddw)[idx], c3));
  op = cadd(op, cscale(ddw, c2));
  op = cadd(op, cscale(dw, c1));
  op = cadd(op, cscale(w, c0));
  double2 v = cscale(op, -dL * or2 * q.visc * q.orho);
  v = cadd(v, cscale(C2(a.s)[idx], a.BuoFac * dL * or2 * a.grav[r]));
  W2(a.impl_w)[idx] = v;
}

static int launch(void (*k)(DcArgs), const DcArgs& a, void* stream) {
  const long n = (long)a.lc.L * a.rc.N;
  if (a.rc.N < 5 || a.lc.L < 1) return MH_EINVAL;
  hipLaunchKernelGGL(k, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, mh_stream(stream), a);
  return mh_launch_status();
}

extern "C" int mh_td_dc(const MhDcArgs* args, void* stream) {
  const DcArgs& a = *args;
  if (!a.Advr || !a.dAdvt_dr || !a.w || !a.dw || !a.ddw || !a.z || !a.dz || !a.ew || !a.ep)
    return MH_EINVAL;
  return launch(k_td_dc, a, stream);
}

extern "C" int mh_buoy_dc(const MhDcArgs* args, void* stream) {
  const DcArgs& a = *args;
  if (!a.rwp || !a.s || !a.ew0 || !a.grav) return MH_EINVAL;
  return launch(k_buoy_dc, a, stream);
}

extern "C" int mh_impl_dc(const MhDcArgs* args, void* stream) {
  const DcArgs& a = *args;
  if (!a.w || !a.dw || !a.ddw || !a.dddw || !a.ddddw || !a.s || !a.impl_w || !a.grav)
    return MH_EINVAL;
  return launch(k_impl_dc, a, stream);
}

__global__ __launch_bounds__(256) void k_affine(const double* __restrict__ A0,
                                                const double* __restrict__ A1, double* out,
                                                long n, double w, const double* wdev,
                                                const int* when) {
  if (when && *when == 0) return;
  if (wdev) w = *wdev;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    out[i] = A0[i] + w * A1[i];
}

extern "C" int mh_affine_f64(const double* A0, const double* A1, double* out, long n, double w,
                             const double* wdev, const int* when, void* stream) {
  if (!A0 || !A1 || !out || n < 0) return MH_EINVAL;
  if (n == 0) return 0;
  long blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_affine, dim3((unsigned)blocks), dim3(256), 0, mh_stream(stream), A0, A1,
                     out, n, w, wdev, when);
  return mh_launch_status();
}
