// Azimuthal real FFTs between Fourier buffers F[c][m][t][b] (complex, m < n_m)
// and grid rows G[c][b][t][phi] (real, n_phi = 2*Mh points), norm="forward"
// (the inverse carries no scale, the forward scales by 1/n_phi).
//
// A real transform of length 2*Mh is done as one complex Stockham FFT of
// length Mh in LDS (radix 2/3/4/5 stages) with the standard even/odd packing.
// One workgroup handles RB consecutive batch rows b of one (component, theta)
// pair, so the F side is read/written in RB-wide contiguous complex runs and
// the G side in whole contiguous rows.
#include "common.h"

typedef MhFftArgs FftArgs;
#define TW(a) reinterpret_cast<const double2*>((a).tw)
#define TW2(a) reinterpret_cast<const double2*>((a).tw2)

template <int R>
__device__ __forceinline__ void dft(double2* v, double dir) {
  if (R == 2) {
    const double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if (R == 3) {
    const double c1 = -0.5, s1 = dir * 0.86602540378443864676372317075294;
    const double2 a = v[0], b = v[1], c = v[2];
    const double2 bc = cadd(b, c), bmc = csub(b, c);
    const double2 t1 = make_double2(a.x + c1 * bc.x, a.y + c1 * bc.y);
    const double2 t2 = make_double2(-s1 * bmc.y, s1 * bmc.x);  // i*s1*(b-c)
    v[0] = cadd(a, bc);
    v[1] = cadd(t1, t2);
    v[2] = csub(t1, t2);
  } else if (R == 4) {
    const double2 a = v[0], b = v[1], c = v[2], d = v[3];
    const double2 apc = cadd(a, c), amc = csub(a, c), bpd = cadd(b, d), bmd = csub(b, d);
    const double2 ibmd = make_double2(-dir * bmd.y, dir * bmd.x);  // i*dir*(b-d)
    v[0] = cadd(apc, bpd);
    v[1] = cadd(amc, ibmd);
    v[2] = csub(apc, bpd);
    v[3] = csub(amc, ibmd);
  } else if (R == 5) {
    const double c1 = 0.30901699437494742410229341718282, c2 = -0.80901699437494742410229341718282;
    const double s1 = dir * 0.95105651629515357211643933337938;
    const double s2 = dir * 0.58778525229247312916870595463907;
    const double2 a = v[0];
    const double2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const double2 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const double2 r1 = make_double2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
    const double2 r2 = make_double2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
    const double2 q1 = make_double2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
    const double2 q2 = make_double2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
    const double2 iq1 = make_double2(-q1.y, q1.x), iq2 = make_double2(-q2.y, q2.x);
    v[0] = make_double2(a.x + t1.x + t2.x, a.y + t1.y + t2.y);
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

// Runs all stages; returns the buffer that holds the result.
__device__ double2* fft_lds(double2* b0, double2* b1, const FftArgs& a, double dir) {
  int Ns = 1;
  double2* in = b0;
  double2* out = b1;
  for (int s = 0; s < a.nstage; ++s) {
    const int R = a.radix[s];
    if (R == 4) stockham_stage<4>(in, out, a.Mh, Ns, a.RB, TW(a), dir);
    else if (R == 2) stockham_stage<2>(in, out, a.Mh, Ns, a.RB, TW(a), dir);
    else if (R == 3) stockham_stage<3>(in, out, a.Mh, Ns, a.RB, TW(a), dir);
    else stockham_stage<5>(in, out, a.Mh, Ns, a.RB, TW(a), dir);
    __syncthreads();
    Ns *= R;
    double2* t = in;
    in = out;
    out = t;
  }
  return in;
}

// F (spectral) -> G (grid): x_j = sum_k X_k e^{+2 pi i jk/n}, X_{n-k} = conj(X_k).
__global__ __launch_bounds__(256) void k_fft_inverse(FftArgs a) {
  extern __shared__ double2 lds[];
  double2* b0 = lds;
  double2* b1 = lds + a.RB * a.Mh;
  const int c = blockIdx.z, t = blockIdx.y, bs = blockIdx.x * a.RB;
  const int Mh = a.Mh;
  const double* Fct = a.F + c * a.Fc_stride + (long)t * 2 * a.FB;
  for (int idx = threadIdx.x; idx < a.RB * Mh; idx += blockDim.x) {
    const int m = idx / a.RB, j = idx - m * a.RB;
    double2 v = make_double2(0.0, 0.0);
    if (m < a.n_m && bs + j < a.B) {
      const double* src = Fct + (long)m * a.Fm_stride + 2L * (bs + j);
      v = make_double2(src[0], m == 0 ? 0.0 : src[1]);
    }
    b0[j * Mh + m] = v;
  }
  __syncthreads();
  // Z_k = E_k + i O_k,  E_k = X_k + conj(X_{Mh-k}),  O_k = (X_k - conj(X_{Mh-k})) e^{+2 pi i k/n}
  for (int idx = threadIdx.x; idx < a.RB * Mh; idx += blockDim.x) {
    const int j = idx / Mh, k = idx - j * Mh;
    const double2 X = b0[j * Mh + k];
    const double2 Xc = (k == 0) ? make_double2(0.0, 0.0) : cconj(b0[j * Mh + Mh - k]);
    const double2 E = cadd(X, Xc);
    const double2 O = cmul(csub(X, Xc), cconj(TW2(a)[k]));
    b1[j * Mh + k] = make_double2(E.x - O.y, E.y + O.x);
  }
  __syncthreads();
  double2* res = fft_lds(b1, b0, a, 1.0);
  for (int idx = threadIdx.x; idx < a.RB * Mh; idx += blockDim.x) {
    const int j = idx / Mh, q = idx - j * Mh;
    if (bs + j < a.B) {
      double2* dst = reinterpret_cast<double2*>(a.G + c * a.Gc_stride +
                                                ((long)(bs + j) * a.T + t) * 2L * Mh);
      dst[q] = res[j * Mh + q];
    }
  }
}

// G (grid) -> F (spectral), X_k = (1/n) sum_j x_j e^{-2 pi i jk/n}, k < n_m,
// optionally times ost2[t] per component.
__global__ __launch_bounds__(256) void k_fft_forward(FftArgs a) {
  extern __shared__ double2 lds[];
  double2* b0 = lds;
  double2* b1 = lds + a.RB * a.Mh;
  const int c = blockIdx.z, t = blockIdx.y, bs = blockIdx.x * a.RB;
  const int Mh = a.Mh;
  for (int idx = threadIdx.x; idx < a.RB * Mh; idx += blockDim.x) {
    const int j = idx / Mh, q = idx - j * Mh;
    double2 v = make_double2(0.0, 0.0);
    if (bs + j < a.B) {
      const double2* src = reinterpret_cast<const double2*>(
          a.G + c * a.Gc_stride + ((long)(bs + j) * a.T + t) * 2L * Mh);
      v = src[q];
    }
    b0[j * Mh + q] = v;
  }
  __syncthreads();
  double2* res = fft_lds(b0, b1, a, -1.0);
  double scale = 1.0 / (2.0 * Mh);
  if (a.comp_scale && a.comp_scale[c]) scale *= a.ost2[t];
  double* Fct = a.F + c * a.Fc_stride + (long)t * 2 * a.FB;
  for (int idx = threadIdx.x; idx < a.RB * a.n_m; idx += blockDim.x) {
    const int m = idx / a.RB, j = idx - m * a.RB;
    if (bs + j >= a.B) continue;
    const double2 Z = res[j * Mh + m];
    const double2 Zc = cconj(res[j * Mh + ((Mh - m) % Mh)]);
    const double2 E = cscale(cadd(Z, Zc), 0.5);
    const double2 D = csub(Z, Zc);
    const double2 O = make_double2(0.5 * D.y, -0.5 * D.x);  // (Z - Zc) / (2i)
    const double2 X = cadd(E, cmul(TW2(a)[m], O));
    double* dst = Fct + (long)m * a.Fm_stride + 2L * (bs + j);
    dst[0] = X.x * scale;
    dst[1] = X.y * scale;
  }
}

extern "C" int mh_fft(int forward, const MhFftArgs* args, void* stream) {
  const FftArgs& a = *args;
  if (a.Mh < 2 || a.RB < 1 || a.B < 1 || a.T < 1 || a.ncomp < 1 || a.n_m > a.Mh) return MH_EINVAL;
  int prod = 1;
  for (int s = 0; s < a.nstage; ++s) {
    const int R = a.radix[s];
    if (R != 2 && R != 3 && R != 4 && R != 5) return MH_EINVAL;
    prod *= R;
  }
  if (prod != a.Mh) return MH_EINVAL;
  const size_t lds = 2ull * a.RB * a.Mh * sizeof(double2);
  if (lds > 160 * 1024) return MH_EINVAL;
  dim3 grid(mh_cdiv(a.B, a.RB), a.T, a.ncomp);
  if (forward)
    hipLaunchKernelGGL(k_fft_forward, grid, dim3(256), lds, mh_stream(stream), a);
  else
    hipLaunchKernelGGL(k_fft_inverse, grid, dim3(256), lds, mh_stream(stream), a);
  return mh_launch_status();
}

