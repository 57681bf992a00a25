// LDS Stockham FFT building blocks (radix 2/3/4/5), shared by the azimuthal
// FFT kernels and the fused grid-space kernel.
#pragma once

#include "common.h"

// In-register DFT of R points, X_k = sum_n x_n exp(dir * 2 pi i nk / R), on
// complex values of type V (double2 or float2; the constants are rounded to
// the element type, so an FP32 transform stays in FP32).
template <int R, typename V>
__device__ __forceinline__ void dft(V* v, double dir) {
  typedef decltype(v[0].x) S;
  if (R == 2) {
    const V a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if (R == 3) {
    const S c1 = (S)-0.5, s1 = (S)(dir * 0.86602540378443864676372317075294);
    const V a = v[0], b = v[1], c = v[2];
    const V bc = cadd(b, c), bmc = csub(b, c);
    const V t1 = mk2(a.x + c1 * bc.x, a.y + c1 * bc.y);
    const V t2 = mk2(-s1 * bmc.y, s1 * bmc.x);  // i*s1*(b-c)
    v[0] = cadd(a, bc);
    v[1] = cadd(t1, t2);
    v[2] = csub(t1, t2);
  } else if (R == 4) {
    const S d = (S)dir;
    const V a = v[0], b = v[1], c = v[2], e = v[3];
    const V apc = cadd(a, c), amc = csub(a, c), bpd = cadd(b, e), bmd = csub(b, e);
    const V ibmd = mk2(-d * bmd.y, d * bmd.x);  // i*dir*(b-d)
    v[0] = cadd(apc, bpd);
    v[1] = cadd(amc, ibmd);
    v[2] = csub(apc, bpd);
    v[3] = csub(amc, ibmd);
  } else if (R == 5) {
    const S c1 = (S)0.30901699437494742410229341718282, c2 = (S)-0.80901699437494742410229341718282;
    const S s1 = (S)(dir * 0.95105651629515357211643933337938);
    const S s2 = (S)(dir * 0.58778525229247312916870595463907);
    const V a = v[0];
    const V t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const V t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const V r1 = mk2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
    const V r2 = mk2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
    const V q1 = mk2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
    const V q2 = mk2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
    const V iq1 = mk2(-q1.y, q1.x), iq2 = mk2(-q2.y, q2.x);
    v[0] = mk2(a.x + t1.x + t2.x, a.y + t1.y + t2.y);
    v[1] = cadd(r1, iq1);
    v[4] = csub(r1, iq1);
    v[2] = cadd(r2, iq2);
    v[3] = csub(r2, iq2);
  }
}

template <int R>
__device__ void stockham_stage(const double2* __restrict__ in, double2* __restrict__ out, int Mh,
                               int Ns, int RB, const double2* __restrict__ tw, double dir) {
  const int nb = Mh / R;
  const int tstep = Mh / (Ns * R);
  for (int idx = threadIdx.x; idx < RB * nb; idx += blockDim.x) {
    const int row = idx / nb, bf = idx - row * nb;
    const int k = bf % Ns;
    const double2* src = in + row * Mh;
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[bf + r * nb];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        double2 w = tw[r * k * tstep];
        w.y *= -dir;  // tw holds exp(-i...); dir=+1 (inverse) conjugates
        v[r] = cmul(v[r], w);
      }
    }
    dft<R>(v, dir);
    double2* dst = out + row * Mh + (bf / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[r * Ns] = v[r];
  }
}

// Runs all stages on RB rows of length Mh; returns the buffer holding the result.
__device__ __forceinline__ double2* fft_lds_rows(double2* b0, double2* b1, int Mh, int nstage,
                                                 const int* radix, int RB,
                                                 const double2* __restrict__ tw, double dir) {
  int Ns = 1;
  double2* in = b0;
  double2* out = b1;
  for (int s = 0; s < nstage; ++s) {
    const int R = radix[s];
    if (R == 4) stockham_stage<4>(in, out, Mh, Ns, RB, tw, dir);
    else if (R == 2) stockham_stage<2>(in, out, Mh, Ns, RB, tw, dir);
    else if (R == 3) stockham_stage<3>(in, out, Mh, Ns, RB, tw, dir);
    else stockham_stage<5>(in, out, Mh, Ns, RB, tw, dir);
    __syncthreads();
    Ns *= R;
    double2* t = in;
    in = out;
    out = t;
  }
  return in;
}

