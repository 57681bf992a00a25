// Spectral-side elementwise kernels of the Boussinesq CNAB2 step: explicit
// terms (nonlinear + Coriolis), IMEX right-hand sides with boundary rows,
// buoyancy coupling and the l=0 pressure right-hand side, implicit terms and
// stored old states, and energy / spectrum reductions.
//
// Every spectral array is (L, N) complex, m-major lm order, N radial points
// per row; the coupled (w, p) right-hand side is (L, 2N).
#include "common.h"

#define C2(p) reinterpret_cast<const double2*>(p)
#define W2(p) reinterpret_cast<double2*>(p)

typedef MhLmCoef LmCoef;

typedef MhRadCoef RadCoef;

// ------------------------------------------------------------ explicit terms
typedef MhTdArgs TdArgs;

__global__ __launch_bounds__(256) void k_td(TdArgs a) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int N = a.rc.N;
  if (idx >= (long)a.lc.L * N) return;
  const int lm = (int)(idx / N), r = (int)(idx - (long)lm * N);
  const int l = a.lc.l[lm], m = a.lc.m[lm];
  const bool bdry = (r == 0) || (r == N - 1);
  const double dLh = a.lc.dLh[lm], dm = (double)m;
  const double or1 = a.rc.or1[r], or2 = a.rc.or2[r], or4 = a.rc.or4[r];
  const double c2 = 2.0 * a.CorFac;
  const bool hasA = l < a.lc.l_max;
  const long iA = hasA ? idx + N : idx;              // (l+1, m)
  const long iS = (l > m) ? idx - N : idx;            // (l-1, m) or self (zero coefficient)
  const double mA = hasA ? 1.0 : 0.0;
  const double2 zA = cscale(C2(a.z)[iA], mA), zS = C2(a.z)[iS];
  const double2 dwA = cscale(C2(a.dw)[iA], mA), dwS = C2(a.dw)[iS];
  const double2 wA = cscale(C2(a.w)[iA], mA), wS = C2(a.w)[iS];
  const double2 dw = C2(a.dw)[idx], w = C2(a.w)[idx], z = C2(a.z)[idx];
  const double t2A = a.lc.dTh2A[lm], t2S = a.lc.dTh2S[lm];
  const double t3A = a.lc.dTh3A[lm], t3S = a.lc.dTh3S[lm];
  const double t4A = a.lc.dTh4A[lm], t4S = a.lc.dTh4S[lm];
  double2 ew = make_double2(0.0, 0.0), ez = ew, ep = ew, eb = ew, ejp = ew;
  if (!bdry) {
    const double2 Advr = C2(a.Advr)[idx];
    double2 cor = cimul(dw, dm);
    cor = cadd(cor, csub(cscale(zA, t2A), cscale(zS, t2S)));
    ew = cadd(cscale(Advr, or2), cscale(cor, c2 * or1));
    double2 ct = cimul(z, dm);
    ct = cadd(ct, cscale(dwA, t3A));
    ct = cadd(ct, cscale(wA, or1 * t4A));
    ct = cadd(ct, cscale(dwS, t3S));
    ct = csub(ct, cscale(wS, or1 * t4S));
    ez = cadd(cscale(C2(a.Advp)[idx], dLh), cscale(ct, c2 * or2));
    double2 cp = cimul(cadd(dw, cscale(w, or1 * dLh)), -dm);
    cp = cadd(cp, cadd(cscale(zA, t3A), cscale(zS, t3S)));
    ep = cadd(cscale(C2(a.Advt)[idx], -dLh), cscale(cp, c2 * or2));
    if (a.l_mag) {
      eb = cscale(C2(a.VxBp)[idx], dLh);
      ejp = cscale(C2(a.VxBr)[idx], dLh * or4);
    }
  }
  W2(a.ew)[idx] = ew;
  W2(a.ez)[idx] = ez;
  W2(a.ep)[idx] = ep;
  // entropy: dLh VSt (zero on the boundaries) - or2 d(dVSr)/dr
  const double2 vs = bdry ? make_double2(0.0, 0.0) : cscale(C2(a.VSt)[idx], dLh);
  W2(a.es)[idx] = csub(vs, cscale(C2(a.dVSr_dr)[idx], or2));
  if (a.l_mag) {
    W2(a.eb)[idx] = eb;
    // induction: dLh or4 VxBr + or2 d(-dLh r^2 VxBt)/dr
    W2(a.ej)[idx] = csub(ejp, cscale(C2(a.dVxBt_dr)[idx], or2 * dLh));
  }
}

extern "C" int mh_td(const MhTdArgs* args, void* stream) {
  const TdArgs& a = *args;
  const long n = (long)a.lc.L * a.rc.N;
  hipLaunchKernelGGL(k_td, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, mh_stream(stream), a);
  return mh_launch_status();
}

// ------------------------------------------------------------ IMEX rhs
typedef MhImexSys ImexSys;

typedef MhRhsArgs RhsArgs;

// weights of the optional terms t4..t7 (BPR353 stages)
struct ExtraW {
  double w[4];
};

__device__ __forceinline__ double2 imex(const ImexSys& s, long i, const RhsArgs& a,
                                        const ExtraW& x) {
  double2 v = cscale(C2(s.old)[i], a.wimp);
  if (s.impl) v = cadd(v, cscale(C2(s.impl)[i], a.wl1));
  if (s.expl0) v = cadd(v, cscale(C2(s.expl0)[i], a.we0));
  if (s.expl1) v = cadd(v, cscale(C2(s.expl1)[i], a.we1));
  if (s.t4) v = cadd(v, cscale(C2(s.t4)[i], x.w[0]));
  if (s.t5) v = cadd(v, cscale(C2(s.t5)[i], x.w[1]));
  if (s.t6) v = cadd(v, cscale(C2(s.t6)[i], x.w[2]));
  if (s.t7) v = cadd(v, cscale(C2(s.t7)[i], x.w[3]));
  return v;
}

__global__ __launch_bounds__(256) void k_rhs(RhsArgs a) {
  ExtraW x = {{0.0, 0.0, 0.0, 0.0}};
  if (a.wdev) {  // weights of this step from device memory (launch arguments fixed)
    a.wimp = a.wdev[0];
    a.wl1 = a.wdev[1];
    a.we0 = a.wdev[2];
    a.we1 = a.wdev[3];
    for (int q = 0; q < 4; ++q) x.w[q] = a.wdev[8 + q];
  }
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int N = a.N;
  if (idx >= (long)a.L * N) return;
  const int lm = (int)(idx / N), r = (int)(idx - (long)lm * N);
  const bool bdry = (r == 0) || (r == N - 1);
  const bool m0 = a.m[lm] == 0;
  auto fix = [&](double2 v) { return m0 ? make_double2(v.x, 0.0) : v; };
  double2 vs = bdry ? make_double2((r == N - 1 && lm == 0) ? a.bots_l0 : 0.0, 0.0)
                    : imex(a.sys[0], idx, a, x);
  W2(a.rs)[idx] = fix(vs);
  W2(a.rz)[idx] = bdry ? make_double2(0.0, 0.0) : fix(imex(a.sys[1], idx, a, x));
  const long iw = (long)lm * 2 * N + r;
  W2(a.rwp)[iw] = bdry ? make_double2(0.0, 0.0) : fix(imex(a.sys[2], idx, a, x));
  W2(a.rwp)[iw + N] = bdry ? make_double2(0.0, 0.0) : fix(imex(a.sys[3], idx, a, x));
  if (a.l_mag) {
    W2(a.rb)[idx] = bdry ? make_double2(0.0, 0.0) : fix(imex(a.sys[4], idx, a, x));
    W2(a.rj)[idx] = bdry ? make_double2(0.0, 0.0) : fix(imex(a.sys[5], idx, a, x));
  }
}

extern "C" int mh_rhs(const MhRhsArgs* args, void* stream) {
  const RhsArgs& a = *args;
  for (int k = 0; k < 6; ++k)
    if ((a.sys[k].t4 || a.sys[k].t5 || a.sys[k].t6 || a.sys[k].t7) && !a.wdev) return MH_EINVAL;
  const long n = (long)a.L * a.N;
  hipLaunchKernelGGL(k_rhs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, mh_stream(stream), a);
  return mh_launch_status();
}

// ------------------------------------------------------------ buoyancy + p0 rhs
typedef MhBuoyArgs BuoyArgs;

__global__ __launch_bounds__(256) void k_buoy(BuoyArgs a) {
  if (a.wdev) a.fac = a.wdev[4];
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int N = a.N;
  if (idx >= (long)a.L * N) return;
  const int lm = (int)(idx / N), r = (int)(idx - (long)lm * N);
  const long iw = (long)lm * 2 * N + r;
  if (r > 0 && r < N - 1) {
    const double2 s = C2(a.s)[idx];
    const double g = a.fac * a.rgrav[r];
    double2 v = W2(a.rwp)[iw];
    v.x += g * s.x;
    v.y += (a.m[lm] == 0) ? 0.0 : g * s.y;
    W2(a.rwp)[iw] = v;
  }
  if (lm == a.lm_l0) {
    // p0 rhs: 0 at the outer boundary, BuoFac rgrav Re(s) + Re(dwdt expl) below
    const double v = (r == 0) ? 0.0 : a.BuoFac * a.rgrav[r] * C2(a.s)[idx].x + C2(a.ew0)[idx].x;
    W2(a.rwp)[iw + N] = make_double2(v, 0.0);
  }
}

extern "C" int mh_buoy(const MhBuoyArgs* args, void* stream) {
  const BuoyArgs& a = *args;
  const long n = (long)a.L * a.N;
  hipLaunchKernelGGL(k_buoy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, mh_stream(stream), a);
  return mh_launch_status();
}

// ------------------------------------------------------------ implicit terms
typedef MhImplArgs ImplArgs;

__global__ __launch_bounds__(256) void k_impl(ImplArgs a) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int N = a.rc.N;
  if (idx >= (long)a.lc.L * N) return;
  const int lm = (int)(idx / N), r = (int)(idx - (long)lm * N);
  const double dLh = a.lc.dLh[lm];
  const double or1 = a.rc.or1[r], or2 = a.rc.or2[r], or3 = a.rc.or3[r];
  const double h = dLh * or2;
  const double2 s = C2(a.s)[idx], z = C2(a.z)[idx], w = C2(a.w)[idx], dw = C2(a.dw)[idx];
  W2(a.old_s)[idx] = s;
  W2(a.impl_s)[idx] = cscale(csub(cadd(C2(a.d2s)[idx], cscale(C2(a.ds)[idx], 2.0 * or1)), cscale(s, h)), a.opr);
  W2(a.old_z)[idx] = cscale(z, h);
  W2(a.impl_z)[idx] = csub(cscale(C2(a.d2z)[idx], h), cscale(z, h * h));
  W2(a.old_w)[idx] = cscale(w, h);
  W2(a.old_p)[idx] = cscale(dw, -h);
  const bool zero = (lm == a.lm_l0) || r == 0 || r == N - 1;
  double2 iw = make_double2(0.0, 0.0), ip = iw;
  if (!zero) {
    iw = cscale(csub(C2(a.ddw)[idx], cscale(w, h)), h);
    iw = csub(iw, C2(a.dp)[idx]);
    iw = cadd(iw, cscale(s, a.BuoFac * a.rc.rgrav[r]));
    double2 t = cscale(C2(a.dddw)[idx], -1.0);
    t = cadd(t, cscale(dw, h));
    t = csub(t, cscale(w, 2.0 * dLh * or3));
    ip = cadd(cscale(C2(a.p)[idx], h), cscale(t, h));
  }
  W2(a.impl_w)[idx] = iw;
  W2(a.impl_p)[idx] = ip;
  if (a.l_mag) {
    const double2 b = C2(a.b)[idx], aj = C2(a.aj)[idx];
    W2(a.old_b)[idx] = cscale(b, h);
    W2(a.old_j)[idx] = cscale(aj, h);
    W2(a.impl_b)[idx] = cscale(csub(C2(a.ddb)[idx], cscale(b, h)), a.opm * h);
    W2(a.impl_j)[idx] = cscale(csub(C2(a.ddj)[idx], cscale(aj, h)), a.opm * h);
  }
}

extern "C" int mh_impl(const MhImplArgs* args, void* stream) {
  const ImplArgs& a = *args;
  const long n = (long)a.lc.L * a.rc.N;
  hipLaunchKernelGGL(k_impl, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, mh_stream(stream), a);
  return mh_launch_status();
}

// ------------------------------------------------------------ energies
// out[q * nout + i] for q = kinetic poloidal, kinetic toroidal, magnetic
// poloidal, magnetic toroidal; i = degree l (by_m = 0) or order m (by_m = 1; zero
// for the orders that are not multiples of minc).
// Radial integration uses the Clenshaw-Curtis weight vector qw.
typedef MhEnergyArgs EnergyArgs;

__device__ __forceinline__ double cc2real(double2 c, bool m0) {
  return m0 ? c.x * c.x : 2.0 * (c.x * c.x + c.y * c.y);
}

__global__ __launch_bounds__(256) void k_energy(EnergyArgs a) {
  const int i = blockIdx.x;  // l or m
  const int N = a.N;
  // by_m 2: per degree, each quantity split into four parts by
  // bin = 2*(m == 0) + ((l + m) odd), out[(4*quantity + bin)*(l_max+1) + l]
  const bool parts = a.by_m == 2;
  double acc[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) acc[q] = 0.0;
  // work items: (m, r) for a fixed l, or (l, r) for a fixed m; the orders are
  // the multiples of minc (mode index m / minc), other orders have no work
  const int minc = a.minc > 1 ? a.minc : 1;
  const int nitem = (a.by_m == 1 ? (i % minc ? 0 : a.l_max - i + 1) : (i / minc + 1)) * N;
  for (int it = threadIdx.x; it < nitem; it += blockDim.x) {
    const int q = it / N, r = it - q * N;
    int l, m;
    if (a.by_m == 1) { m = i; l = i + q; } else { l = i; m = q * minc; }
    const long lm = a.lm_start[m / minc] + (l - m);
    const long idx = lm * N + r;
    const bool m0 = m == 0;
    const double dLh = (double)l * (double)(l + 1);
    const double wq = a.qw[r], or2 = a.or2[r];
    const double wk = a.orho1 ? wq * a.orho1[r] : wq;  // kinetic: 1/rho0 (anelastic)
    if (!parts) {
      acc[0] += wk * 0.5 * dLh * (dLh * or2 * cc2real(C2(a.w)[idx], m0) + cc2real(C2(a.dw)[idx], m0));
      acc[1] += wk * 0.5 * dLh * cc2real(C2(a.z)[idx], m0);
      if (a.l_mag) {
        acc[2] += wq * 0.5 * a.LFfac * dLh * (dLh * or2 * cc2real(C2(a.b)[idx], m0) + cc2real(C2(a.db)[idx], m0));
        acc[3] += wq * 0.5 * a.LFfac * dLh * cc2real(C2(a.aj)[idx], m0);
      }
    } else {
      double e[4];
      e[0] = wk * 0.5 * dLh * (dLh * or2 * cc2real(C2(a.w)[idx], m0) + cc2real(C2(a.dw)[idx], m0));
      e[1] = wk * 0.5 * dLh * cc2real(C2(a.z)[idx], m0);
      e[2] = e[3] = 0.0;
      if (a.l_mag) {
        e[2] = wq * 0.5 * a.LFfac * dLh * (dLh * or2 * cc2real(C2(a.b)[idx], m0) + cc2real(C2(a.db)[idx], m0));
        e[3] = wq * 0.5 * a.LFfac * dLh * cc2real(C2(a.aj)[idx], m0);
      }
      const int bin = 2 * (m0 ? 1 : 0) + ((l + m) & 1);
      // selects, not an indexed register array (which would go to scratch)
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) acc[4 * k + bb] += bb == bin ? e[k] : 0.0;
    }
  }
  __shared__ double sh[4][256];
  const int npass = parts ? 4 : 1;
  for (int pass = 0; pass < npass; ++pass) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      double v = acc[q];
#pragma unroll
      for (int pp = 1; pp < 4; ++pp)
        if (pass == pp) v = acc[4 * pp + q];
      sh[q][threadIdx.x] = v;
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s)
        for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0)
      for (int q = 0; q < 4; ++q) a.out[(4 * pass + q) * (a.l_max + 1) + i] = sh[q][0];
    __syncthreads();  // the next pass overwrites sh
  }
}

// by_m 3: radial profiles per degree, summed over the orders of each degree
// and not integrated, for the length scales of get_dlm (output.py:750-861):
// out[(q*(l_max+1) + l)*N + r] for q = kinetic poloidal, toroidal, the same
// two over m > 0 only, then the four magnetic ones (times LFfac).  One
// workgroup per degree, consecutive lanes on consecutive radial levels.
__global__ __launch_bounds__(128) void k_energy_lr(EnergyArgs a) {
  const int l = blockIdx.x;
  const int N = a.N;
  const double dLh = (double)l * (double)(l + 1);
  const long plane = (long)(a.l_max + 1) * N;
  const int minc = a.minc > 1 ? a.minc : 1;
  for (int r = threadIdx.x; r < N; r += blockDim.x) {
    const double or2 = a.or2[r];
    const double ok = a.orho1 ? a.orho1[r] : 1.0;  // kinetic: 1/rho0 (anelastic)
    double e[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) e[q] = 0.0;
    for (int m = 0; m <= l; m += minc) {
      const long idx = (long)(a.lm_start[m / minc] + (l - m)) * N + r;
      const bool m0 = m == 0;
      const double kp = ok * 0.5 * dLh * (dLh * or2 * cc2real(C2(a.w)[idx], m0) + cc2real(C2(a.dw)[idx], m0));
      const double kt = ok * 0.5 * dLh * cc2real(C2(a.z)[idx], m0);
      e[0] += kp;
      e[1] += kt;
      e[2] += m0 ? 0.0 : kp;
      e[3] += m0 ? 0.0 : kt;
      if (a.l_mag) {
        const double mp = 0.5 * a.LFfac * dLh * (dLh * or2 * cc2real(C2(a.b)[idx], m0) + cc2real(C2(a.db)[idx], m0));
        const double mt = 0.5 * a.LFfac * dLh * cc2real(C2(a.aj)[idx], m0);
        e[4] += mp;
        e[5] += mt;
        e[6] += m0 ? 0.0 : mp;
        e[7] += m0 ? 0.0 : mt;
      }
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) a.out[q * plane + (long)l * N + r] = e[q];
  }
}

extern "C" int mh_energy(const MhEnergyArgs* args, void* stream) {
  const EnergyArgs& a = *args;
  if (a.by_m < 0 || a.by_m > 3 || a.l_max < 0 || a.N < 1 || a.minc < 0 || a.out == nullptr ||
      a.w == nullptr || a.dw == nullptr || a.z == nullptr || a.lm_start == nullptr ||
      a.or2 == nullptr || (a.by_m != 3 && a.qw == nullptr) ||
      (a.l_mag && (a.b == nullptr || a.db == nullptr || a.aj == nullptr)))
    return MH_EINVAL;
  if (a.by_m == 3)
    hipLaunchKernelGGL(k_energy_lr, dim3(a.l_max + 1), dim3(128), 0, mh_stream(stream), a);
  else
    hipLaunchKernelGGL(k_energy, dim3(a.l_max + 1), dim3(256), 0, mh_stream(stream), a);
  return mh_launch_status();
}

