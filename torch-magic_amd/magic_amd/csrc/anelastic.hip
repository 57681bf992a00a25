// Anelastic (polytropic reference state) step: the grid products with the
// Courant maxima, the azimuthal derivative in Fourier space, the entropy
// heating and 1/rho0 finish, the anelastic implicit terms, the
// angular-momentum corrections and the integral condition of the l=0
// pressure.  The transforms, solves and explicit terms are the Boussinesq
// path's kernels; the operators are assembled on the host.
#include "common.h"

typedef MhNlAnelArgs NlAnelArgs;
typedef MhImplAnelArgs ImplAnelArgs;
typedef MhImplArgs ImplArgs;

#define C2(p) reinterpret_cast<const double2*>(p)
#define W2(p) reinterpret_cast<double2*>(p)

namespace {

__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// One workgroup row per radial level (blockIdx.y), grid points strided over
// blockIdx.x; velocity-derived inputs read as zero on the boundary levels.
__global__ __launch_bounds__(256) void k_nl_anel(NlAnelArgs a) {
  const int r = blockIdx.y;
  const long npt = (long)a.T * a.P;
  const long lvl = (long)r * npt;
  const long N = a.N;
  const long cv = a.Gv_c_stride, cb = a.Gb_c_stride, cd = a.Gd_c_stride;
  const double or1 = a.or1[r], or2 = a.or2[r], or4 = a.or4[r];
  const double rr = 1.0 / or1;
  const double beta = a.prof.beta[r], orho1 = a.prof.orho1[r], otemp1 = a.prof.otemp1[r];
  const double visc = a.prof.visc[r];
  const bool bulk = (r > 0) && (r < a.N - 1);
  const double prefac = a.prof.ViscHeatFac * or4 * orho1 * otemp1 * visc;
  const double ohm = a.prof.OhmLossFac * or2 * otemp1;  // lambda = 1
  double mr = 0.0, mh = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < npt;
       i += (long)gridDim.x * blockDim.x) {
    const int t = (int)(i / a.P);
    const double ost2 = a.ost2[t], cosn = a.cosn2[t];
    const long o = lvl + i;
    double br = 0.0, bt = 0.0, bp = 0.0;
    if (a.l_mag) {
      br = a.Gb[o];
      bt = a.Gb[cb + o];
      bp = a.Gb[2 * cb + o];
    }
    double vr = 0.0, vt = 0.0, vp = 0.0;
    if (bulk) {
      vr = a.Gv[o];
      vt = a.Gv[cv + o];
      vp = a.Gv[2 * cv + o];
      const double dvrdr = a.Gv[N * npt + o], dvtdr = a.Gv[cv + N * npt + o],
                   dvpdr = a.Gv[2 * cv + N * npt + o];
      const double dvrdt = a.Gv[cv + 2 * N * npt + o], dvrdp = a.Gv[2 * cv + 2 * N * npt + o];
      const double dvtdp = a.Gd[o] * ost2, dvpdp = a.Gd[cd + o] * ost2;
      const double cvr = a.Gs[N * npt + o];
      const double sc = a.Gs[o];
      double Advr = -or2 * orho1 * (vr * (dvrdr - (2.0 * or1 + beta) * vr) +
                                    ost2 * (vt * (dvrdt - rr * vt) + vp * (dvrdp - rr * vp)));
      double Advt = or4 * orho1 * (-vr * (dvtdr - beta * vt) + vt * (cosn * vt + dvpdp + dvrdr) +
                                   vp * (cosn * vp - dvtdp));
      double Advp = or4 * orho1 * (-vr * (dvpdr - beta * vp) - vt * (dvtdp + cvr) - vp * dvpdp);
      double VxBr = 0.0, VxBt = 0.0, VxBp = 0.0;
      const double br_ = beta * rr;
      const double e_rr = dvrdr - (2.0 * or1 + beta) * vr;
      const double e_tt = cosn * vt + dvpdp + dvrdr - or1 * vr;
      const double e_pp = dvpdp + cosn * vt + or1 * vr;
      const double e_tp = 2.0 * dvtdp + cvr - 2.0 * cosn * vp;
      const double e_rt = rr * dvtdr - (2.0 + br_) * vt + or1 * dvrdt;
      const double e_rp = rr * dvpdr - (2.0 + br_) * vp + or1 * dvrdp;
      double heat = prefac * (2.0 * e_rr * e_rr + 2.0 * e_tt * e_tt + 2.0 * e_pp * e_pp +
                              e_tp * e_tp + ost2 * (e_rt * e_rt + e_rp * e_rp) -
                              2.0 / 3.0 * (beta * vr) * (beta * vr));
      if (a.l_mag) {
        const double cbr = a.Gb[N * npt + o], cbt = a.Gb[cb + N * npt + o],
                     cbp = a.Gb[2 * cb + N * npt + o];
        Advr += a.LFfac * ost2 * (cbt * bp - cbp * bt);
        Advt += a.LFfac * or4 * (cbp * br - cbr * bp);
        Advp += a.LFfac * or4 * (cbr * bt - cbt * br);
        VxBr = orho1 * ost2 * (vt * bp - vp * bt);
        VxBt = orho1 * or4 * (vp * br - vr * bp);
        VxBp = orho1 * or4 * (vr * bt - vt * br);
        if (a.prof.OhmLossFac != 0.0)
          heat += ohm * (or2 * cbr * cbr + ost2 * (cbt * cbt + cbp * cbp));
      }
      const long oo = (long)(r - 1) * npt + i;
      const long cs = a.Go_c_stride;
      a.Go[oo] = Advr;
      a.Go[cs + oo] = vr * sc;
      a.Go[2 * cs + oo] = VxBr;
      a.Go[3 * cs + oo] = heat;
      a.Go[4 * cs + oo] = Advt;
      a.Go[5 * cs + oo] = Advp;
      a.Go[6 * cs + oo] = or2 * vt * sc;
      a.Go[7 * cs + oo] = or2 * vp * sc;
      a.Go[8 * cs + oo] = VxBt;
      a.Go[9 * cs + oo] = VxBp;
    }
    const double o2 = orho1 * orho1;
    const double vflr2 = o2 * vr * vr;
    const double vflh2 = (vt * vt + vp * vp) * ost2 * o2;
    double cr, ch;
    if (a.l_mag) {
      const double valr = br * br * a.LFfac * orho1;
      const double valh2 = (bt * bt + bp * bp) * a.LFfac * ost2 * orho1;
      double valr2 = valr, valh2m = valh2;
      if (a.alf_damp) {
        valr2 = valr * valr / (valr + a.valri2[r]);
        valh2m = valh2 * valh2 / (valh2 + a.valhi2[r]);
      }
      cr = or4 * (a.cf2 * vflr2 + a.af2 * valr2);
      ch = or2 * (a.cf2 * vflh2 + a.af2 * valh2m);
    } else {
      cr = a.cf2 * or4 * vflr2;
      ch = a.cf2 * or2 * vflh2;
    }
    mr = fmax(mr, cr);
    mh = fmax(mh, ch);
  }
  __shared__ double sr[4], sh[4];
  mr = wave_max(mr);
  mh = wave_max(mh);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sr[w] = mr;
    sh[w] = mh;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double vr_ = fmax(fmax(sr[0], sr[1]), fmax(sr[2], sr[3]));
    const double vh_ = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    atomicMax(a.cfl + r, (unsigned long long)__double_as_longlong(vr_));
    atomicMax(a.cfl + a.N + r, (unsigned long long)__double_as_longlong(vh_));
  }
}

__global__ void k_zero_cfl(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// out[k][m][t][b] = i (m minc) F[c0 + k][m][t][b]
__global__ __launch_bounds__(256) void k_fourier_dphi(MhDphiArgs a) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)a.n_m * a.T * a.B;
  if (idx >= 2 * per) return;
  const int k = (int)(idx / per);
  long rem = idx - k * per;
  const int m = (int)(rem / ((long)a.T * a.B));
  rem -= (long)m * a.T * a.B;
  const int t = (int)(rem / a.B), b = (int)(rem - (long)t * a.B);
  const double om = (double)m * (double)(a.minc > 1 ? a.minc : 1);
  const double2 v = C2(a.F + (a.c0 + k) * a.Fc_stride + m * a.Fm_stride + (long)t * 2 * a.FB)[b];
  W2(a.out + k * a.Oc_stride + m * a.Om_stride + (long)t * 2 * a.OB)[b] =
      make_double2(-om * v.y, om * v.x);
}

__global__ __launch_bounds__(256) void k_entropy_anel(MhEntropyAnelArgs a) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)a.L * a.N) return;
  const int r = (int)(idx % a.N);
  const double f = a.orho1[r];
  const double2 e = C2(a.es)[idx], h = C2(a.heat)[idx];
  W2(a.es)[idx] = make_double2(f * (e.x + h.x), f * (e.y + h.y));
}

__global__ __launch_bounds__(256) void k_impl_anel(ImplAnelArgs A) {
  const ImplArgs& a = A.base;
  const MhAnelProf& pr = A.prof;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int N = a.rc.N;
  if (idx >= (long)a.lc.L * N) return;
  const int lm = (int)(idx / N), r = (int)(idx - (long)lm * N);
  const double dLh = a.lc.dLh[lm];
  const double or1 = a.rc.or1[r], or2 = a.rc.or2[r];
  const double h = dLh * or2;
  const double beta = pr.beta[r], dbeta = pr.dbeta[r], visc = pr.visc[r], dLv = pr.dLvisc[r];
  const double2 s = C2(a.s)[idx], z = C2(a.z)[idx], w = C2(a.w)[idx], dw = C2(a.dw)[idx];
  const double2 ddw = C2(a.ddw)[idx], p = C2(a.p)[idx];
  W2(a.old_s)[idx] = s;
  const double d1s = beta + pr.dLtemp0[r] + 2.0 * or1 + pr.dLkappa[r];
  W2(a.impl_s)[idx] = cscale(csub(cadd(C2(a.d2s)[idx], cscale(C2(a.ds)[idx], d1s)), cscale(s, h)),
                             a.opr * pr.kappa[r]);
  W2(a.old_z)[idx] = cscale(z, h);
  const double z0 = dLv * beta + 2.0 * dLv * or1 + dbeta + 2.0 * beta * or1;
  const double2 dz = C2(A.dz)[idx];
  W2(a.impl_z)[idx] = cscale(csub(cadd(C2(a.d2z)[idx], cscale(dz, dLv - beta)), cscale(z, z0 + h)),
                             h * visc);
  W2(a.old_w)[idx] = cscale(w, h);
  W2(a.old_p)[idx] = cscale(dw, -h);
  double2 iw = make_double2(0.0, 0.0), ip = iw;
  const bool zero = A.zero_edges && ((lm == a.lm_l0) || r == 0 || r == N - 1);
  if (!zero) {
    const double cw = h + 4.0 / 3.0 * (dbeta + dLv * beta + (3.0 * dLv + beta) * or1);
    double2 dif = csub(cadd(ddw, cscale(dw, 2.0 * dLv - beta / 3.0)), cscale(w, cw));
    dif = cscale(dif, h * visc);
    iw = cadd(csub(cscale(p, beta), C2(a.dp)[idx]), dif);
    iw = cadd(iw, cscale(s, a.BuoFac * pr.rgrav_rho[r]));
    double2 t = cscale(C2(a.dddw)[idx], -1.0);
    t = cadd(t, cscale(ddw, beta - dLv));
    t = cadd(t, cscale(dw, h + dLv * beta + dbeta + 2.0 * (dLv + beta) * or1));
    t = csub(t, cscale(w, h * (2.0 * or1 + 2.0 / 3.0 * beta + dLv)));
    ip = cadd(cscale(p, h), cscale(t, visc * h));
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

// block sum of three values over 256 threads
__device__ void block_sum3(double& x, double& y, double& z) {
  __shared__ double sh[3][256];
  sh[0][threadIdx.x] = x;
  sh[1][threadIdx.x] = y;
  sh[2][threadIdx.x] = z;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
      sh[2][threadIdx.x] += sh[2][threadIdx.x + s];
    }
    __syncthreads();
  }
  x = sh[0][0];
  y = sh[1][0];
  z = sh[2][0];
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_am_correct(MhAmArgs a) {
  const int N = a.N;
  double sz = 0.0, sx = 0.0, sy = 0.0;
  for (int r = threadIdx.x; r < N; r += blockDim.x) {
    if (a.do_z) sz += C2(a.z)[(long)a.l1m0 * N + r].x * a.r2q[r];
    if (a.do_e && a.l1m1 >= 0) {
      const double2 v = C2(a.z)[(long)a.l1m1 * N + r];
      sx += v.x * a.r2q[r];
      sy += v.y * a.r2q[r];
    }
  }
  block_sum3(sz, sx, sy);
  const double cz = a.cz * sz;
  const double2 ce = make_double2(a.ce * sx, a.ce * sy);
  for (int r = threadIdx.x; r < N; r += blockDim.x) {
    const double p0 = a.pz[r], p1 = a.pdz[r], p2 = a.pd2z[r];
    if (a.do_z) {
      const long i = (long)a.l1m0 * N + r;
      W2(a.z)[i].x -= p0 * cz;
      W2(a.dz)[i].x -= p1 * cz;
      W2(a.d2z)[i].x -= p2 * cz;
    }
    if (a.do_e && a.l1m1 >= 0) {
      const long i = (long)a.l1m1 * N + r;
      W2(a.z)[i] = csub(C2(a.z)[i], cscale(ce, p0));
      W2(a.dz)[i] = csub(C2(a.dz)[i], cscale(ce, p1));
      W2(a.d2z)[i] = csub(C2(a.d2z)[i], cscale(ce, p2));
    }
  }
}

__global__ __launch_bounds__(256) void k_p0_integral(MhP0IntArgs a) {
  const int N = a.N;
  double v = 0.0, d1 = 0.0, d2 = 0.0;
  for (int r = threadIdx.x; r < N; r += blockDim.x)
    v += a.u[r] * C2(a.s)[(long)a.lm_l0 * N + r].x;
  block_sum3(v, d1, d2);
  if (threadIdx.x == 0) W2(a.rwp)[(long)a.lm_l0 * 2 * N + N] = make_double2(v, 0.0);
}

}  // namespace

extern "C" int mh_nl_anel(const MhNlAnelArgs* args, void* stream) {
  const NlAnelArgs& a = *args;
  if (a.N < 3 || a.T < 1 || a.P < 1 || !a.Gs || !a.Gv || !a.Gd || !a.Go || !a.cfl ||
      (a.l_mag && !a.Gb) || !a.prof.beta || !a.prof.orho1 || !a.prof.otemp1 || !a.prof.visc)
    return MH_EINVAL;
  hipLaunchKernelGGL(k_zero_cfl, dim3(mh_cdiv(2 * a.N, 256)), dim3(256), 0, mh_stream(stream),
                     a.cfl, 2 * a.N);
  const long npt = (long)a.T * a.P;
  int gx = (int)((npt + 1023) / 1024);
  if (gx > 1024) gx = 1024;
  hipLaunchKernelGGL(k_nl_anel, dim3(gx, a.N), dim3(256), 0, mh_stream(stream), a);
  return mh_launch_status();
}

extern "C" int mh_fourier_dphi(const MhDphiArgs* args, void* stream) {
  const MhDphiArgs& a = *args;
  if (!a.F || !a.out || a.B < 1 || a.B > a.FB || a.B > a.OB || a.n_m < 1 || a.T < 1 || a.c0 < 0)
    return MH_EINVAL;
  const long n = 2L * a.n_m * a.T * a.B;
  hipLaunchKernelGGL(k_fourier_dphi, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     mh_stream(stream), a);
  return mh_launch_status();
}

extern "C" int mh_entropy_anel(const MhEntropyAnelArgs* args, void* stream) {
  const MhEntropyAnelArgs& a = *args;
  if (!a.es || !a.heat || !a.orho1 || a.L < 1 || a.N < 2) return MH_EINVAL;
  const long n = (long)a.L * a.N;
  hipLaunchKernelGGL(k_entropy_anel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     mh_stream(stream), a);
  return mh_launch_status();
}

extern "C" int mh_impl_anel(const MhImplAnelArgs* args, void* stream) {
  const ImplAnelArgs& A = *args;
  const MhAnelProf& p = A.prof;
  if (!p.beta || !p.dbeta || !p.visc || !p.dLvisc || !p.kappa || !p.dLkappa || !p.dLtemp0 ||
      !p.rgrav_rho || !A.dz)
    return MH_EINVAL;
  const long n = (long)A.base.lc.L * A.base.rc.N;
  hipLaunchKernelGGL(k_impl_anel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     mh_stream(stream), A);
  return mh_launch_status();
}

extern "C" int mh_am_correct(const MhAmArgs* args, void* stream) {
  const MhAmArgs& a = *args;
  if (!a.z || !a.dz || !a.d2z || !a.r2q || !a.pz || !a.pdz || !a.pd2z || a.N < 2 || a.l1m0 < 0)
    return MH_EINVAL;
  hipLaunchKernelGGL(k_am_correct, dim3(1), dim3(256), 0, mh_stream(stream), a);
  return mh_launch_status();
}

extern "C" int mh_p0_integral(const MhP0IntArgs* args, void* stream) {
  const MhP0IntArgs& a = *args;
  if (!a.rwp || !a.s || !a.u || a.N < 2 || a.lm_l0 < 0) return MH_EINVAL;
  hipLaunchKernelGGL(k_p0_integral, dim3(1), dim3(256), 0, mh_stream(stream), a);
  return mh_launch_status();
}
