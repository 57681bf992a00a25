// Implicit-operator rebuild on the device (runs when dt changes) and banded
// LINPACK-order factor/solve for the finite-difference per-degree systems.
//
//  * k_build_*: per-degree collocation matrices of the Boussinesq Chebyshev
//    implicit systems (entropy, toroidal, poloidal/pressure 2N, magnetic b and
//    j, l=0 pressure), with the boundary-coefficient factor 1/2 on the first
//    and last Chebyshev columns.
//  * k_gj_inverse: one workgroup per matrix, row (and optionally column)
//    max-norm equilibration, Gauss-Jordan inversion with partial pivoting
//    (first maximum), result rescaled so that inv @ rhs solves the original
//    system.  info[b] = first zero pivot (1-based) or 0.
//  * k_gemm_f64: batched C = alpha * A B with general strides on FP64 MFMA,
//    used to fold the DCT-I and the radial derivative matrices into the
//    inverses.
#include <mutex>

#include "common.h"

typedef MhBuildArgs BuildArgs;

__device__ __forceinline__ double bfac(int n, int N) { return (n == 0 || n == N - 1) ? 0.5 : 1.0; }

// Entry (i, j) of the collocation matrix of system `kind` at degree l.
__device__ double build_entry(const BuildArgs& a, int kind, int l, int i, int j) {
  const int N = a.N;
  const double dL = (double)l * (double)(l + 1);
  const double rn = a.rnorm, w0 = a.wimp_dev ? *a.wimp_dev : a.wimp;
  double v = 0.0;
  if (kind <= 3 || kind == 5) {
    const int n = j;
    const double R = a.R[i * N + n], D = a.D1[i * N + n], D2 = a.D2[i * N + n];
    const double or1 = a.or1[i], or2 = a.or2[i];
    const bool top = (i == 0), bot = (i == N - 1);
    if (kind == 0) {
      v = rn * (R - w0 * a.opr * (D2 + 2.0 * or1 * D - dL * or2 * R));
      if (top || bot) v = rn * R;
    } else if (kind == 1) {
      if (l == 0) {
        v = (i == j) ? 1.0 : 0.0;
      } else {
        v = rn * dL * or2 * (R - w0 * (D2 - dL * or2 * R));
        if (top) v = (a.ktopv == 1) ? rn * (D - 2.0 * or1 * R) : rn * R;
        if (bot) v = (a.kbotv == 1) ? rn * (D - 2.0 * or1 * R) : rn * R;
      }
    } else if (kind == 2 || kind == 3) {
      if (l == 0) {
        v = (i == j) ? 1.0 : 0.0;
      } else {
        v = rn * dL * or2 * (R - w0 * a.opm * (D2 - dL * or2 * R));
        if (kind == 2) {
          if (top) v = rn * (D + (double)l * or1 * R);
          if (bot) v = rn * (D - (double)(l + 1) * or1 * R);
        } else if (top || bot) {
          v = rn * R;
        }
      }
    } else {  // p0
      v = top ? rn * R : rn * D;
    }
    const bool ident = (kind == 1 || kind == 2 || kind == 3) && l == 0;
    if (!ident) {
      if ((top || (bot && kind != 5)) && n >= a.n_cheb) v = 0.0;
      v *= bfac(n, N);
    }
  } else {  // WP, 2N x 2N
    const int bi = i / N, ii = i - bi * N;  // row block, row within block
    const int bj = j / N, n = j - bj * N;   // column block, Chebyshev index
    if (l == 0) {
      v = (i == j) ? 1.0 : 0.0;
    } else {
      const double R = a.R[ii * N + n], D = a.D1[ii * N + n], D2 = a.D2[ii * N + n],
                   D3 = a.D3[ii * N + n];
      const double or1 = a.or1[ii], or2 = a.or2[ii];
      const bool top = (ii == 0), bot = (ii == N - 1);
      if (!top && !bot) {
        if (bi == 0 && bj == 0) v = rn * dL * or2 * (R - w0 * (D2 - dL * or2 * R));
        else if (bi == 0 && bj == 1) v = rn * w0 * D;
        else if (bi == 1 && bj == 0)
          v = rn * dL * or2 * (-D - w0 * (-D3 + dL * or2 * D - dL * or2 * 2.0 * or1 * R));
        else v = -rn * w0 * dL * or2 * R;
      } else if (bj == 0) {
        if (bi == 0) {
          v = rn * R;  // w = 0 on both boundaries
        } else {
          const int ks = top ? a.ktopv : a.kbotv;
          v = (ks == 1) ? rn * (D2 - 2.0 * or1 * D) : rn * D;
        }
      }
      if ((top || bot) && n >= a.n_cheb) v = 0.0;
      v *= bfac(n, N);
    }
  }
  return v;
}

// kind 0..5: one system into a.out; kind 6: systems 0..nsys-1 stacked in
// a.out ([k][l][i][j]), the (w,p) system in a.out_wp and p0 in a.out_p0.
__global__ void k_build(BuildArgs a) {
  if (a.when && *a.when == 0) return;
  const int N = a.N, lp1 = a.l_max + 1;
  const long nn = (long)N * N, ww = 4L * nn;
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int kind = a.kind, l, i, j;
  double* out;
  if (kind != 6) {
    const int n2 = (kind == 4) ? 2 * N : N;
    const long per = (long)n2 * n2;
    if (idx >= per * ((kind == 5) ? 1 : lp1)) return;
    l = (int)(idx / per);
    const int rem = (int)(idx - (long)l * per);
    i = rem / n2;
    j = rem - i * n2;
    out = a.out + idx;
  } else {
    const long n_stack = (long)a.nsys * lp1 * nn;
    if (idx < n_stack) {
      const long b = idx / nn;
      kind = (int)(b / lp1);
      l = (int)(b - (long)kind * lp1);
      const int rem = (int)(idx - b * nn);
      i = rem / N;
      j = rem - i * N;
      out = a.out + idx;
    } else if ((idx -= n_stack) < lp1 * ww) {
      kind = 4;
      l = (int)(idx / ww);
      const int rem = (int)(idx - (long)l * ww);
      i = rem / (2 * N);
      j = rem - i * (2 * N);
      out = a.out_wp + idx;
    } else if ((idx -= lp1 * ww) < nn) {
      kind = 5;
      l = 0;
      i = (int)(idx / N);
      j = (int)(idx - (long)i * N);
      out = a.out_p0 + idx;
    } else {
      return;
    }
  }
  *out = build_entry(a, kind, l, i, j);
}

extern "C" int mh_build_matrices(const MhBuildArgs* args, void* stream) {
  const BuildArgs& a = *args;
  if (a.N < 3 || a.kind < 0 || a.kind > 6) return MH_EINVAL;
  if (a.kind == 6 && (a.nsys < 1 || a.nsys > 4 || !a.out_wp || !a.out_p0)) return MH_EINVAL;
  const long nn = (long)a.N * a.N, lp1 = a.l_max + 1;
  long total;
  if (a.kind == 6) total = a.nsys * lp1 * nn + lp1 * 4 * nn + nn;
  else if (a.kind == 4) total = lp1 * 4 * nn;
  else total = nn * ((a.kind == 5) ? 1 : lp1);
  hipLaunchKernelGGL(k_build, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, mh_stream(stream), a);
  return mh_launch_status();
}

// ------------------------------------------------------------ Gauss-Jordan
// work: per matrix n x 2n doubles; A and out may alias.
__global__ __launch_bounds__(256) void k_gj_inverse(const double* __restrict__ A, double* out,
                                                    double* __restrict__ work,
                                                    double* __restrict__ scal, int n,
                                                    int precond, int* info, const int* when) {
  if (when && *when == 0) return;
  const int bm = blockIdx.x;
  const long nn = (long)n * n;
  const double* Am = A + bm * nn;
  double* W = work + bm * 2 * nn;
  double* fr = scal + (long)bm * 2 * n;
  double* fc = fr + n;
  const int n2 = 2 * n;
  __shared__ double sval[256];
  __shared__ int sidx[256];
  __shared__ int s_piv;
  __shared__ double s_pv;
  __shared__ int s_info;
  if (threadIdx.x == 0) s_info = 0;
  // row equilibration
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    double mx = 0.0;
    for (int j = 0; j < n; ++j) mx = fmax(mx, fabs(Am[(long)i * n + j]));
    fr[i] = (precond >= 1 && mx > 0.0) ? 1.0 / mx : 1.0;
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    double mx = 0.0;
    for (int i = 0; i < n; ++i) mx = fmax(mx, fabs(fr[i] * Am[(long)i * n + j]));
    fc[j] = (precond >= 2 && mx > 0.0) ? 1.0 / mx : 1.0;
  }
  __syncthreads();
  for (long e = threadIdx.x; e < (long)n * n2; e += blockDim.x) {
    const int i = (int)(e / n2), j = (int)(e - (long)i * n2);
    W[e] = (j < n) ? fr[i] * Am[(long)i * n + j] * fc[j] : ((j - n) == i ? 1.0 : 0.0);
  }
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    // pivot: first maximum of |W[i][k]|, i >= k
    double best = -1.0;
    int bi = k;
    for (int i = k + threadIdx.x; i < n; i += blockDim.x) {
      const double v = fabs(W[(long)i * n2 + k]);
      if (v > best) { best = v; bi = i; }
    }
    sval[threadIdx.x] = best;
    sidx[threadIdx.x] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) {
        const double o = sval[threadIdx.x + s];
        const int oi = sidx[threadIdx.x + s];
        if (o > sval[threadIdx.x] || (o == sval[threadIdx.x] && oi < sidx[threadIdx.x])) {
          sval[threadIdx.x] = o;
          sidx[threadIdx.x] = oi;
        }
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      s_piv = sidx[0];
      s_pv = W[(long)sidx[0] * n2 + k];
      if (fabs(s_pv) <= 1e-15 && s_info == 0) s_info = k + 1;
    }
    __syncthreads();
    const int p = s_piv;
    const double pinv = (s_pv != 0.0) ? 1.0 / s_pv : 0.0;
    if (p != k) {
      for (int j = threadIdx.x; j < n2; j += blockDim.x) {
        const double t = W[(long)k * n2 + j];
        W[(long)k * n2 + j] = W[(long)p * n2 + j];
        W[(long)p * n2 + j] = t;
      }
      __syncthreads();
    }
    for (int j = threadIdx.x; j < n2; j += blockDim.x) W[(long)k * n2 + j] *= pinv;
    __syncthreads();
    // eliminate column k from all other rows (columns k+1 .. 2n-1 change)
    const long ncol = n2 - k - 1;
    for (long e = threadIdx.x; e < (long)n * ncol; e += blockDim.x) {
      const int i = (int)(e / ncol);
      const int j = k + 1 + (int)(e - (long)i * ncol);
      if (i == k) continue;
      W[(long)i * n2 + j] -= W[(long)i * n2 + k] * W[(long)k * n2 + j];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x)
      if (i != k) W[(long)i * n2 + k] = 0.0;
    __syncthreads();
  }
  double* O = out + bm * nn;
  for (long e = threadIdx.x; e < nn; e += blockDim.x) {
    const int i = (int)(e / n), j = (int)(e - (long)i * n);
    O[e] = fc[i] * W[(long)i * n2 + n + j] * fr[j];
  }
  if (threadIdx.x == 0 && info) info[bm] = s_info;
}


// ------------------------------------------------- register-resident Gauss-Jordan
// One workgroup of TR x TC threads holds the whole (equilibrated) matrix in
// registers: thread (tr, tc) owns rows tr + TR*u (u < RPT) and columns
// tc + TC*v (v < CPT).  In-place Gauss-Jordan with implicit partial pivoting
// (first maximum of |a[i][k]| over the rows not yet used): at step k the
// pivot column and the normalised pivot row are broadcast through LDS and
// every thread updates its block; column k takes the overlay of the identity.
// With r_k the pivot row of step k and pi(r_k) = k, the array ends holding
// storage[s][c] = inv[pi(s)][r_c], which the final store un-permutes.  Two
// barriers per step (LDS broadcast buffers alternate between steps).
template <int TR, int TC, int RPT, int CPT>
__global__ __launch_bounds__(TR * TC) void k_gj_reg(const double* A, double* out, int n,
                                                    int precond, int* info, const int* when) {
  if (when && *when == 0) return;
  constexpr int RMAX = TR * RPT, CMAX = TC * CPT;
  __shared__ double colbuf[2][RMAX];
  __shared__ double rowbuf[2][CMAX];
  __shared__ double cval[2][TR];
  __shared__ int cidx[2][TR];
  __shared__ int piv_row[CMAX];
  __shared__ int step_of[RMAX];
  __shared__ double s_fr[RMAX];
  __shared__ unsigned long long s_fc[CMAX];
  const int tid = threadIdx.x, tr = tid / TC, tc = tid - tr * TC;
  const long nn = (long)n * n;
  const double* Am = A + (long)blockIdx.x * nn;
  double a[RPT][CPT];
#pragma unroll
  for (int u = 0; u < RPT; ++u)
#pragma unroll
    for (int v = 0; v < CPT; ++v) {
      const int i = tr + TR * u, j = tc + TC * v;
      a[u][v] = (i < n && j < n) ? Am[(long)i * n + j] : 0.0;
    }
  for (int j = tid; j < CMAX; j += TR * TC) s_fc[j] = 0ull;
  // row equilibration: the TC threads of a row group are consecutive lanes
  double fr[RPT];
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    double mx = 0.0;
#pragma unroll
    for (int v = 0; v < CPT; ++v) mx = fmax(mx, fabs(a[u][v]));
    for (int o = TC / 2; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    fr[u] = (precond >= 1 && mx > 0.0) ? 1.0 / mx : 1.0;
#pragma unroll
    for (int v = 0; v < CPT; ++v) a[u][v] *= fr[u];
    const int i = tr + TR * u;
    if (tc == 0 && i < RMAX) s_fr[i] = fr[u];
  }
  __syncthreads();
  // column equilibration (max over the TR row groups through LDS)
#pragma unroll
  for (int v = 0; v < CPT; ++v) {
    double mx = 0.0;
#pragma unroll
    for (int u = 0; u < RPT; ++u) mx = fmax(mx, fabs(a[u][v]));
    atomicMax(&s_fc[tc + TC * v], (unsigned long long)__double_as_longlong(mx));
  }
  __syncthreads();
  double fc[CPT];
#pragma unroll
  for (int v = 0; v < CPT; ++v) {
    const double mx = __longlong_as_double((long long)s_fc[tc + TC * v]);
    fc[v] = (precond >= 2 && mx > 0.0) ? 1.0 / mx : 1.0;
#pragma unroll
    for (int u = 0; u < RPT; ++u) a[u][v] *= fc[v];
  }
  __syncthreads();  // all reads of s_fc done before it is rewritten below
#pragma unroll
  for (int v = 0; v < CPT; ++v)
    if (tr == 0) s_fc[tc + TC * v] = (unsigned long long)__double_as_longlong(fc[v]);
  unsigned used = 0u;  // bit u: row tr + TR*u has been a pivot row
  int first_zero = 0;
  for (int k = 0; k < n; ++k) {
    const int buf = k & 1;
    const int kc = k % TC, kv = k / TC;
    // (a) holders of column k: candidates, column values, clear column k
    if (tc == kc) {
      double best = -1.0;
      int bi = 0x7fffffff;
#pragma unroll
      for (int u = 0; u < RPT; ++u) {
        const int i = tr + TR * u;
        double x = 0.0;
#pragma unroll
        for (int v = 0; v < CPT; ++v)
          if (v == kv) x = a[u][v];
        if (i < RMAX) colbuf[buf][i] = x;
        if (i < n && !((used >> u) & 1u) && fabs(x) > best) {
          best = fabs(x);
          bi = i;
        }
#pragma unroll
        for (int v = 0; v < CPT; ++v)
          if (v == kv) a[u][v] = 0.0;
      }
      cval[buf][tr] = best;
      cidx[buf][tr] = bi;
    }
    __syncthreads();
    // (b) pivot (same decision in every thread), normalised pivot row
    double best = -1.0;
    int r = 0x7fffffff;
#pragma unroll
    for (int q = 0; q < TR; ++q) {
      const double cv = cval[buf][q];
      const int ci = cidx[buf][q];
      if (cv > best || (cv == best && ci < r)) {
        best = cv;
        r = ci;
      }
    }
    double pv = colbuf[buf][r];
    if (!(fabs(pv) > 1e-15)) {
      if (first_zero == 0) first_zero = k + 1;
      if (pv == 0.0) pv = 1.0;
    }
    const double pinv = 1.0 / pv;
    const int rtr = r % TR, ru = r / TR;
    if (tr == rtr) {
#pragma unroll
      for (int u = 0; u < RPT; ++u)
        if (u == ru) {
          used |= 1u << u;
#pragma unroll
          for (int v = 0; v < CPT; ++v) {
            const int j = tc + TC * v;
            rowbuf[buf][j] = (j == k) ? pinv : a[u][v] * pinv;
          }
        }
    }
    if (tid == 0) {
      piv_row[k] = r;
      step_of[r] = k;
    }
    __syncthreads();
    // (c) elimination
    double rw[CPT];
#pragma unroll
    for (int v = 0; v < CPT; ++v) rw[v] = rowbuf[buf][tc + TC * v];
#pragma unroll
    for (int u = 0; u < RPT; ++u) {
      const int i = tr + TR * u;
      if (i == r) {
#pragma unroll
        for (int v = 0; v < CPT; ++v) a[u][v] = rw[v];
      } else {
        const double f = (i < RMAX) ? colbuf[buf][i] : 0.0;
#pragma unroll
        for (int v = 0; v < CPT; ++v) a[u][v] -= f * rw[v];
      }
    }
  }
  __syncthreads();
  // out[pi(s)][r_c] = fc[pi(s)] * storage[s][c] * fr[r_c] (the equilibration of
  // row pi(s) and column r_c of the original system)
  double* O = out + (long)blockIdx.x * nn;
#pragma unroll
  for (int u = 0; u < RPT; ++u) {
    const int s = tr + TR * u;
    if (s >= n) continue;
    const int oi = step_of[s];
    const double fci = __longlong_as_double((long long)s_fc[oi]);
#pragma unroll
    for (int v = 0; v < CPT; ++v) {
      const int c = tc + TC * v;
      if (c >= n) continue;
      const int oj = piv_row[c];
      O[(long)oi * n + oj] = fci * a[u][v] * s_fr[oj];
    }
  }
  if (tid == 0 && info) info[blockIdx.x] = first_zero;
}

extern "C" int mh_gj_inverse(const double* A, double* out, double* work, double* scal, int n,
                             int nbatch, int precond, int* info, const int* when, void* stream) {
  if (n < 1 || nbatch < 1) return MH_EINVAL;
  if (n <= 112)
    hipLaunchKernelGGL((k_gj_reg<16, 16, 7, 7>), dim3(nbatch), dim3(256), 0, mh_stream(stream), A,
                       out, n, precond, info, when);
  else if (n <= 208)
    hipLaunchKernelGGL((k_gj_reg<16, 32, 13, 7>), dim3(nbatch), dim3(512), 0, mh_stream(stream), A,
                       out, n, precond, info, when);
  else
    hipLaunchKernelGGL(k_gj_inverse, dim3(nbatch), dim3(256), 0, mh_stream(stream), A, out, work,
                       scal, n, precond, info, when);
  return mh_launch_status();
}

// ------------------------------------------- Gauss-Jordan without allocation
// The generic kernel's work space (n x 2n matrix and 2n scale factors per
// system) taken from a file-scope device array instead of caller buffers.
// Host code never passes `scratch` itself to a kernel: its name is a host-side
// stand-in.  The device address comes from hipGetSymbolAddress, looked up once
// per device with that device current and kept in g_scratch_dev.
static __device__ __align__(256) double scratch[MH_GJ_SCRATCH_DOUBLES];

static std::mutex g_scratch_mu;
static double* g_scratch_dev[MH_MAX_DEVICES];

static int gj_scratch_ptr(double** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MH_MAX_DEVICES) {
    (void)hipGetLastError();
    return MH_ELAUNCH;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[dev] == nullptr) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || p == nullptr) {
      (void)hipGetLastError();
      return MH_ELAUNCH;
    }
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_dev[dev];
  return MH_OK;
}

extern "C" int mh_gj_inverse_noalloc(const double* A, double* out, int n, int nbatch,
                                     int precond, int* info, const int* when, void* stream) {
  if (n < 1 || nbatch < 1) return MH_EINVAL;
  if (n <= 208) return mh_gj_inverse(A, out, nullptr, nullptr, n, nbatch, precond, info, when, stream);
  // work: nbatch * 2 n^2 doubles, then the scale factors: nbatch * 2 n
  const long per = 2L * n * n + 2L * n;
  if ((long)nbatch * per > (long)MH_GJ_SCRATCH_DOUBLES) return MH_ETOOBIG;
  double* base = nullptr;
  const int rc = gj_scratch_ptr(&base);
  if (rc != MH_OK) return rc;
  double* work = base;
  double* scal = base + (long)nbatch * 2 * n * n;
  hipLaunchKernelGGL(k_gj_inverse, dim3(nbatch), dim3(256), 0, mh_stream(stream), A, out, work,
                     scal, n, precond, info, when);
  return mh_launch_status();
}

// ------------------------------------------------------------ batched GEMM
typedef MhGemmArgs GemmArgs;

// one wave per 16 x 64 tile of C
__global__ __launch_bounds__(64) void k_gemm_f64(GemmArgs g) {
  if (g.when && *g.when == 0) return;
  const int lane = threadIdx.x;
  const int b = blockIdx.z;
  const int i0 = blockIdx.y * 16, j0 = blockIdx.x * 64;
  const int kk = lane >> 4, c16 = lane & 15;
  const double* A = g.A + b * g.sAb;
  const double* B = g.B + b * g.sBb;
  double4_t acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = (double4_t){0.0, 0.0, 0.0, 0.0};
  const int ia = i0 + c16;
  for (int k0 = 0; k0 < g.K; k0 += 4) {
    const int k = k0 + kk;
    const double av = (ia < g.M && k < g.K) ? A[ia * g.sAr + k * g.sAc] : 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int jb = j0 + 16 * q + c16;
      const double bv = (jb < g.Nn && k < g.K) ? B[k * g.sBr + jb * g.sBc] : 0.0;
      acc[q] = mfma_f64(av, bv, acc[q]);
    }
  }
  double* C = g.C + b * g.sCb;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int ic = i0 + mfma_row(lane, i);
    if (ic >= g.M) continue;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int jc = j0 + 16 * q + c16;
      if (jc < g.Nn) C[ic * g.sCr + jc * g.sCc] = g.alpha * acc[q][i];
    }
  }
}

extern "C" int mh_gemm_f64(const MhGemmArgs* args, void* stream) {
  const GemmArgs& g = *args;
  if (g.M < 1 || g.Nn < 1 || g.K < 1 || g.batch < 1) return MH_EINVAL;
  dim3 grid(mh_cdiv(g.Nn, 64), mh_cdiv(g.M, 16), g.batch);
  hipLaunchKernelGGL(k_gemm_f64, grid, dim3(64), 0, mh_stream(stream), g);
  return mh_launch_status();
}

