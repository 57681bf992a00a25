// Grid-space nonlinear products and per-level Courant maxima (Boussinesq).
//
// Inputs (FP64 grid rows, [comp][batch][t][phi]):
//   s grid   Gs[r]                        (N levels)
//   vectors  Gv[c][g*N + r], c = r/theta/phi component, g = 0 B, 1 curl B,
//            2 u, 3 curl u (sin(theta)-weighted components as synthesised)
// Outputs (bulk levels r = 1..N-2 only, [comp][r-1][t][phi]):
//   0 Advr 1 VSr 2 VxBr 3 Advt 4 Advp 5 VSt 6 VSp 7 VxBt 8 VxBp
// Courant: cfl[r] = max over the level of or4 (cf2 vr^2 + af2 valr2),
//          cfl[N + r] = max of or2 (cf2 vh^2 ost2 + af2 valh2m), stored as the
//          bit pattern of a non-negative double (atomicMax on uint64).
#include "common.h"

typedef MhNlArgs NlArgs;

__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

__global__ __launch_bounds__(256) void k_get_nl(NlArgs a) {
  const int r = blockIdx.y;
  const long npt = (long)a.T * a.P;
  const long lvl = (long)r * npt;
  const long ncv = a.Gv_c_stride;
  const long N = a.N;
  const double or2 = a.or2[r], or4 = a.or4[r];
  const bool bulk = (r > 0) && (r < a.N - 1);
  double mr = 0.0, mh = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < npt;
       i += (long)gridDim.x * blockDim.x) {
    const int t = (int)(i / a.P);
    const double ost2 = a.ost2[t];
    const long o = lvl + i;
    const double vr = a.Gv[(2 * N) * npt + o], vt = a.Gv[ncv + (2 * N) * npt + o],
                 vp = a.Gv[2 * ncv + (2 * N) * npt + o];
    double br = 0.0, bt = 0.0, bp = 0.0;
    if (a.l_mag) {
      br = a.Gv[o];
      bt = a.Gv[ncv + o];
      bp = a.Gv[2 * ncv + o];
    }
    if (bulk) {
      const double cvr = a.Gv[(3 * N) * npt + o], cvt = a.Gv[ncv + (3 * N) * npt + o],
                   cvp = a.Gv[2 * ncv + (3 * N) * npt + o];
      const double sc = a.Gs[o];
      double Advr = -ost2 * (cvt * vp - cvp * vt);
      double Advt = -or4 * (cvp * vr - cvr * vp);
      double Advp = -or4 * (cvr * vt - cvt * vr);
      double VxBr = 0.0, VxBt = 0.0, VxBp = 0.0;
      if (a.l_mag) {
        const double cbr = a.Gv[N * npt + o], cbt = a.Gv[ncv + N * npt + o],
                     cbp = a.Gv[2 * ncv + N * npt + o];
        Advr += a.LFfac * ost2 * (cbt * bp - cbp * bt);
        Advt += a.LFfac * or4 * (cbp * br - cbr * bp);
        Advp += a.LFfac * or4 * (cbr * bt - cbt * br);
        VxBr = ost2 * (vt * bp - vp * bt);
        VxBt = or4 * (vp * br - vr * bp);
        VxBp = or4 * (vr * bt - vt * br);
      }
      const long oo = (long)(r - 1) * npt + i;
      const long cs = a.Go_c_stride;
      a.Go[oo] = Advr;
      a.Go[cs + oo] = vr * sc;
      a.Go[2 * cs + oo] = VxBr;
      a.Go[3 * cs + oo] = Advt;
      a.Go[4 * cs + oo] = Advp;
      a.Go[5 * cs + oo] = or2 * vt * sc;
      a.Go[6 * cs + oo] = or2 * vp * sc;
      a.Go[7 * cs + oo] = VxBt;
      a.Go[8 * cs + oo] = VxBp;
    }
    const double vflr2 = vr * vr;
    const double vflh2 = (vt * vt + vp * vp) * ost2;
    double cr, ch;
    if (a.l_mag) {
      const double valr = br * br * a.LFfac;
      const double valh2 = (bt * bt + bp * bp) * a.LFfac * ost2;
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
    double vr_ = fmax(fmax(sr[0], sr[1]), fmax(sr[2], sr[3]));
    double vh_ = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    atomicMax(a.cfl + r, (unsigned long long)__double_as_longlong(vr_));
    atomicMax(a.cfl + a.N + r, (unsigned long long)__double_as_longlong(vh_));
  }
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

extern "C" int mh_get_nl(const MhNlArgs* args, void* stream) {
  const NlArgs& a = *args;
  if (a.N < 3 || a.T < 1 || a.P < 1) return MH_EINVAL;
  hipLaunchKernelGGL(k_zero_u64, dim3(mh_cdiv(2 * a.N, 256)), dim3(256), 0, mh_stream(stream), a.cfl, 2 * a.N);
  const long npt = (long)a.T * a.P;
  int gx = (int)((npt + 1023) / 1024);
  if (gx > 1024) gx = 1024;
  hipLaunchKernelGGL(k_get_nl, dim3(gx, a.N), dim3(256), 0, mh_stream(stream), a);
  return mh_launch_status();
}

