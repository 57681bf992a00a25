// Azimuthal real FFTs between Fourier buffers F[c][m][t][b] (complex, m < n_m)
// and grid rows G[c][b][t][phi] (real, n_phi = 2*Mh points), norm="forward"
// (the inverse carries no scale, the forward scales by 1/n_phi).
//
// A real transform of length 2*Mh is done as one complex Stockham FFT of
// length Mh in LDS (radix 2/3/4/5 stages) with the standard even/odd packing.
// One workgroup handles RB consecutive batch rows b of one (component, theta)
// pair, so the F side is read/written in RB-wide contiguous complex runs and
// the G side in whole contiguous rows.
#include "fft_lds.h"

typedef MhFftArgs FftArgs;
#define TW(a) reinterpret_cast<const double2*>((a).tw)
#define TW2(a) reinterpret_cast<const double2*>((a).tw2)

__device__ __forceinline__ double2* fft_lds(double2* b0, double2* b1, const FftArgs& a, double dir) {
  return fft_lds_rows(b0, b1, a.Mh, a.nstage, a.radix, a.RB, TW(a), dir);
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

