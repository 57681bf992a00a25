// Legendre tables and MFMA Legendre synthesis / analysis (FP64, or FP32 with
// FP32 accumulation) for gfx950.
//
// Table layouts (per azimuthal order m and parity p = (l-m)%2, l = m+p+2k):
//   synthesis  syn[m][p][k][j]  (k < K4 = roundup(K,4), j < Hp)  j fastest
//   analysis   ana[m][p][j][k]  (j < Hp, k < K16 = roundup(K,16)) k fastest
// j indexes the northern Gauss nodes (theta pairs), Hp = roundup(H,32).
// mp_info[(m*2+p)*4 + {0,1,2,3}] = {K, syn_off, ana_off, K16}.
//
// Fourier buffers F[c][m][t][b] (complex) hold, per component c and order m,
// the (theta, batch) matrix; t = 2j is the northern and t = 2j+1 the southern
// node of pair j (interleaved rows, as in the grid layout).
#include "common.h"

// ---------------------------------------------------------------- P_lm tables
// Orthonormal P_lm without Condon-Shortley phase and sin(theta) dP_lm/dtheta by
// the three-term recurrence (the same recurrences as the reference plms.py).
__global__ void k_plm_tables(int l_max, int H, int Hp, const double* __restrict__ cos_t,
                             const double* __restrict__ sin_t, const double* __restrict__ wgt,
                             const int* __restrict__ mp_info, double* __restrict__ synP,
                             double* __restrict__ synD, double* __restrict__ anaW,
                             double* __restrict__ anaWd) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  if (j >= H) return;
  const double c = cos_t[j], s = sin_t[j];
  const double w2pi = 2.0 * M_PI * wgt[j];
  const double dnorm = 0.28209479177387814347403972578;  // 1/sqrt(4 pi)
  double fac = 1.0;
  for (int q = 3; q <= 2 * m + 1; q += 2) fac = fac * (double)q / (double)(q - 1);
  double pmm = sqrt(fac);
  if (s != 0.0) {
    double sp = 1.0;
    for (int q = 0; q < m; ++q) sp *= s;
    pmm *= sp;
  } else if (m != 0) {
    pmm = 0.0;
  }
  double pm1 = 0.0, p0 = pmm;  // P_{l-1}, P_l (unnormalised)
  for (int l = m; l <= l_max; ++l) {
    const int L1 = l + 1;  // P_{l+1}
    double p1;
    if ((L1 - m) * (L1 + m) > 0) {
      const double dL = (double)L1, dM = (double)m;
      const double a = sqrt((2.0 * dL - 1.0) * (2.0 * dL + 1.0) / ((dL - dM) * (dL + dM)));
      const double b = sqrt((2.0 * dL + 1.0) * (dL + dM - 1.0) * (dL - dM - 1.0) /
                            ((2.0 * dL - 3.0) * (dL - dM) * (dL + dM)));
      p1 = c * a * p0 - b * pm1;
    } else {
      p1 = 0.0;
    }
    const double dl = (double)l, dm = (double)m;
    const double cl1 = sqrt((dl + dm + 1.0) * (dl - dm + 1.0) / ((2.0 * dl + 1.0) * (2.0 * dl + 3.0)));
    const double cl = (l > m) ? sqrt((dl + dm) * (dl - dm) / ((2.0 * dl - 1.0) * (2.0 * dl + 1.0))) : 0.0;
    const double P = dnorm * p0;
    const double dP = dnorm * ((double)l * cl1 * p1 - (double)(l + 1) * cl * pm1);
    const int p = (l - m) & 1, k = (l - m) >> 1;
    const int* info = mp_info + (m * 2 + p) * 4;
    const long so = (long)info[1] + (long)k * Hp + j;
    synP[so] = P;
    synD[so] = dP;
    const long ao = (long)info[2] + (long)j * info[3] + k;
    anaW[ao] = w2pi * P;
    anaWd[ao] = w2pi * dP;
    pm1 = p0;
    p0 = p1;
  }
}

extern "C" int mh_plm_tables(int l_max, int n_m, int H, int Hp, const double* cos_t,
                             const double* sin_t, const double* wgt, const int* mp_info,
                             double* synP, double* synD, double* anaW, double* anaWd,
                             void* stream) {
  if (l_max < 0 || n_m != l_max + 1 || H <= 0 || Hp < H) return MH_EINVAL;
  dim3 grid(mh_cdiv(H, 64), n_m);
  hipLaunchKernelGGL(k_plm_tables, grid, dim3(64), 0, mh_stream(stream), l_max, H, Hp, cos_t, sin_t, wgt,
                     mp_info, synP, synD, anaW, anaWd);
  return mh_launch_status();
}

// ---------------------------------------------------------------- synthesis
typedef MhSynIn SynIn;
typedef MhSynArgs SynArgs;

__device__ __forceinline__ double syn_load(const SynIn& in, long lm, int l, int b, int reim,
                                           int ldx, const double* or2) {
  const long off = lm * 2 * (long)ldx + 2 * b + reim;
  const double dl = (double)l * (double)(l + 1);
  if (in.mode == 0) return in.x1[off];
  if (in.mode == 1) return dl * in.x1[off];
  return -in.x1[off] + or2[b] * dl * in.x2[off];
}

// i*m*z where lane pairs (2q, 2q+1) hold (re, im) of z
__device__ __forceinline__ double imul_pair(double v, double m, int lane) {
  const double o = __shfl_xor(v, 1);
  return (lane & 1) ? m * o : -m * o;
}

// Offset (in doubles) of the real part of F(c, m, t, b), layout F[c][m][t][b].
__device__ __forceinline__ long fidx(long cs, long ms, int B, int c, int m, int t, int b) {
  return c * cs + m * ms + (long)t * 2 * B + 2L * b;
}

// Element type of the Legendre contraction: FP64 (16x16x4 f64 MFMA) or FP32
// (16x16x4 f32 MFMA, FP32 accumulation; tables stored in FP32, inputs rounded
// to FP32 as they are loaded).
template <typename T> struct Mfma;
template <> struct Mfma<double> {
  typedef double4_t acc_t;
  static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) { return mfma_f64(a, b, c); }
  // row of the 16x16 accumulator held in element v of a lane
  static __device__ __forceinline__ int row(int lane, int v) { return mfma_row(lane, v); }
};
template <> struct Mfma<float> {
  typedef float4_t acc_t;
  static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) { return mfma_f32(a, b, c); }
  static __device__ __forceinline__ int row(int lane, int v) { return mfma_row_f32(lane, v); }
};

// WG = 4 waves; wave w covers theta pairs j0 + 16*(w&1) .. and complex
// columns b0 + 8*(w>>1) .. of the groups laid side by side (column bc is
// batch column bc % B of group bc / B, which is also its column in F);
// WG tile = 32 theta pairs x 16 complex columns.
__device__ __forceinline__ SynIn syn_in(const SynArgs& a, int g, int c) {
  // per-lane group select without dynamic indexing of the argument struct
  SynIn r = a.in[0][c];
  if (g == 1) r = a.in[1][c];
  if (g == 2) r = a.in[2][c];
  if (g == 3) r = a.in[3][c];
  return r;
}

template <bool VEC, typename T>
__global__ __launch_bounds__(256, 3) void k_syn(SynArgs a, int n_m) {
  typedef Mfma<T> MF;
  typedef typename MF::acc_t acc_t;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int b0 = blockIdx.x * 16 + 8 * (w >> 1);
  const int j0 = blockIdx.y * 32 + 16 * (w & 1);
  const int col = lane & 15, kk = lane >> 4;
  const int bc = b0 + (col >> 1), reim = col & 1;
  const bool bval = bc < a.ngroups * a.B;
  const int g = bval ? bc / a.B : 0;
  const int bl = bval ? bc - g * a.B : 0;
  const SynIn in[3] = {syn_in(a, g, 0), syn_in(a, g, 1), syn_in(a, g, 2)};
  {
    const int m = blockIdx.z;
    const long lms = a.lm_start[m];
    acc_t acc[VEC ? 10 : 2];
#pragma unroll
    for (int q = 0; q < (VEC ? 10 : 2); ++q) acc[q] = (acc_t){0, 0, 0, 0};
    // Operands of one k-step (4 degrees); the next step's are loaded before
    // the current step's MFMAs are issued, so their latency is hidden.
    struct Ops {
      T aP, aD, xq, xs, xt;
    };
#pragma unroll
    for (int p = 0; p < 2; ++p) {
      const int* info = a.mp_info + (m * 2 + p) * 4;
      const int K = info[0];
      const T* tP = reinterpret_cast<const T*>(a.synP) + info[1];
      const T* tD = reinterpret_cast<const T*>(a.synD) + info[1];
      auto load = [&](int k0) {
        Ops o;
        const int k = k0 + kk;
        const bool v = bval && (k < K);
        const int l = m + p + 2 * k;
        const long lm = lms + p + 2 * k;
        o.aP = tP[(long)k * a.Hp + j0 + col];
        o.xq = v ? (T)syn_load(in[0], lm, l, bl, reim, a.ldx, a.or2) : (T)0;
        if (VEC) {
          o.aD = tD[(long)k * a.Hp + j0 + col];
          o.xs = v ? (T)syn_load(in[1], lm, l, bl, reim, a.ldx, a.or2) : (T)0;
          o.xt = v ? (T)syn_load(in[2], lm, l, bl, reim, a.ldx, a.or2) : (T)0;
        } else {
          o.aD = o.xs = o.xt = (T)0;
        }
        return o;
      };
      if (K <= 0) continue;
      Ops cur = load(0);
      for (int k0 = 0; k0 < K; k0 += 4) {
        Ops nxt = cur;
        if (k0 + 4 < K) nxt = load(k0 + 4);
        if (!VEC) {
          acc[p] = MF::run(cur.aP, cur.xq, acc[p]);
        } else {
          acc[5 * p + 0] = MF::run(cur.aP, cur.xq, acc[5 * p + 0]);  // P Q
          acc[5 * p + 1] = MF::run(cur.aP, cur.xt, acc[5 * p + 1]);  // P T
          acc[5 * p + 2] = MF::run(cur.aP, cur.xs, acc[5 * p + 2]);  // P S
          acc[5 * p + 3] = MF::run(cur.aD, cur.xs, acc[5 * p + 3]);  // dP S
          acc[5 * p + 4] = MF::run(cur.aD, cur.xt, acc[5 * p + 4]);  // dP T
        }
        cur = nxt;
      }
    }
    const double dm = (double)m;
    const int bo = bval ? bc : 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int j = j0 + MF::row(lane, i);
      const bool st = bval && (j < a.H);
      const long oN = fidx(a.c_stride, a.m_stride, a.Bout, 0, m, 2 * j, bo) + reim;
      const long oS = fidx(a.c_stride, a.m_stride, a.Bout, 0, m, 2 * j + 1, bo) + reim;
      if (!VEC) {
        const double e = acc[0][i], o = acc[1][i];
        if (st) {
          a.F[oN] = e + o;
          a.F[oS] = e - o;
        }
      } else {
        const double Qe = acc[0][i], Qo = acc[5][i];
        const double PTe = acc[1][i], PTo = acc[6][i];
        const double PSe = acc[2][i], PSo = acc[7][i];
        const double DSe = acc[3][i], DSo = acc[8][i];
        const double DTe = acc[4][i], DTo = acc[9][i];
        const double iPTp = imul_pair(PTe + PTo, dm, lane), iPTm = imul_pair(PTe - PTo, dm, lane);
        const double iPSp = imul_pair(PSe + PSo, dm, lane), iPSm = imul_pair(PSe - PSo, dm, lane);
        if (st) {
          a.F[oN] = Qe + Qo;
          a.F[oS] = Qe - Qo;
          a.F[a.c_stride + oN] = (DSe + DSo) + iPTp;
          a.F[a.c_stride + oS] = -(DSe - DSo) + iPTm;
          a.F[2 * a.c_stride + oN] = iPSp - (DTe + DTo);
          a.F[2 * a.c_stride + oS] = iPSm + (DTe - DTo);
        }
      }
    }
  }
}

extern "C" int mh_leg_synth(int vector, const MhSynArgs* args, int n_m, void* stream) {
  const SynArgs& a = *args;
  if (a.ngroups < 1 || a.ngroups > 4 || a.B < 1 || a.H < 1 || a.Hp % 32 != 0 || a.Hp < a.H ||
      a.Bout != a.ngroups * a.B)
    return MH_EINVAL;
  if (a.prec != 0 && a.prec != 1) return MH_EINVAL;
  dim3 grid(mh_cdiv(a.ngroups * a.B, 16), mh_cdiv(a.H, 32), n_m);
  hipStream_t st = mh_stream(stream);
  if (a.prec == 0) {
    if (vector) hipLaunchKernelGGL((k_syn<true, double>), grid, dim3(256), 0, st, a, n_m);
    else hipLaunchKernelGGL((k_syn<false, double>), grid, dim3(256), 0, st, a, n_m);
  } else {
    if (vector) hipLaunchKernelGGL((k_syn<true, float>), grid, dim3(256), 0, st, a, n_m);
    else hipLaunchKernelGGL((k_syn<false, float>), grid, dim3(256), 0, st, a, n_m);
  }
  return mh_launch_status();
}


// ---------------------------------------------------------------- analysis
typedef MhAnaArgs AnaArgs;

// WG = 4 waves: wave w covers k0 + 16*(w&1) .. (16 rows of l) and complex batch
// columns b0 + 8*(w>>1) ..; K loop runs over theta pairs in steps of 4.
template <bool VEC, typename T>
__global__ __launch_bounds__(256, 6) void k_ana(AnaArgs a, int n_m) {
  typedef Mfma<T> MF;
  typedef typename MF::acc_t acc_t;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int* tl = a.tiles + 3 * blockIdx.y;
  const int p = tl[1];
  const int kbase = tl[2] + 16 * (w & 1);
  const int nbt = (a.B + 15) / 16;
  const int comp = blockIdx.x / nbt;
  const int b0 = (blockIdx.x % nbt) * 16 + 8 * (w >> 1);
  const int col = lane & 15, kk = lane >> 4;
  const int b = b0 + (col >> 1), reim = col & 1;
  const bool bval = b < a.B;
  const int bl = bval ? b : 0;
  const double sg = p ? -1.0 : 1.0;  // same-parity fold: N + sg*S
  const int arow = kbase + col;      // row of A handled by this lane
  {
    const int m = tl[0];
    if (m >= n_m) return;
    const int* info = a.mp_info + (m * 2 + p) * 4;
    const int K = info[0], K16 = info[3];
    const T* tW = reinterpret_cast<const T*>(a.anaW) + info[2];
    const T* tD = reinterpret_cast<const T*>(a.anaWd) + info[2];
    acc_t acc[VEC ? 4 : 1];
#pragma unroll
    for (int q = 0; q < (VEC ? 4 : 1); ++q) acc[q] = (acc_t){0, 0, 0, 0};
    if (kbase < K) {
      // Operands of one theta step (4 pairs); the next step's are loaded before
      // the current step's MFMAs are issued.
      struct Ops {
        T aw, ad, x1, x2, x3, x4;
      };
      const int c1 = a.cf1[comp], c2 = VEC ? a.cf2[comp] : 0;
      auto load = [&](int j0) {
        Ops o;
        const int j = j0 + kk;
        o.aw = tW[(long)j * K16 + arow];
        o.ad = VEC ? tD[(long)j * K16 + arow] : (T)0;
        o.x1 = o.x2 = o.x3 = o.x4 = (T)0;
        if (bval && j < a.H) {
          const long iN = fidx(a.c_stride, a.m_stride, a.Bin, 0, m, 2 * j, bl) + reim;
          const long iS = fidx(a.c_stride, a.m_stride, a.Bin, 0, m, 2 * j + 1, bl) + reim;
          const double f1N = a.F[c1 * a.c_stride + iN], f1S = a.F[c1 * a.c_stride + iS];
          o.x1 = (T)(f1N + sg * f1S);  // same fold
          if (VEC) {
            const double f2N = a.F[c2 * a.c_stride + iN], f2S = a.F[c2 * a.c_stride + iS];
            o.x2 = (T)(f1N - sg * f1S);  // opposite fold
            o.x3 = (T)(f2N + sg * f2S);
            o.x4 = (T)(f2N - sg * f2S);
          }
        }
        return o;
      };
      Ops cur = load(0);
      for (int j0 = 0; j0 < a.H; j0 += 4) {
        Ops nxt = cur;
        if (j0 + 4 < a.H) nxt = load(j0 + 4);
        if (!VEC) {
          acc[0] = MF::run(cur.aw, cur.x1, acc[0]);
        } else {
          acc[0] = MF::run(cur.ad, cur.x4, acc[0]);  // wdP f2(opposite fold)
          acc[1] = MF::run(cur.ad, cur.x2, acc[1]);  // wdP f1(opposite fold)
          acc[2] = MF::run(cur.aw, cur.x1, acc[2]);  // wP f1(same fold)
          acc[3] = MF::run(cur.aw, cur.x3, acc[3]);  // wP f2(same fold)
        }
        cur = nxt;
      }
    }
    const double dm = (double)m;
    const long lms = a.lm_start[m];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = kbase + MF::row(lane, i);
      const int l = m + p + 2 * k;
      const bool st = bval && (k < K);
      const long off = (lms + p + 2 * k) * 2L * a.ldo + 2 * bl + reim;
      if (!VEC) {
        if (st) a.outS[comp][off] = (double)acc[0][i];
      } else {
        const double inv = (l > 0) ? 1.0 / ((double)l * (double)(l + 1)) : 0.0;
        const double Sv = (double)acc[0][i] + imul_pair((double)acc[2][i], -dm, lane);
        const double Tv = -(double)acc[1][i] + imul_pair((double)acc[3][i], -dm, lane);
        if (st) {
          a.outS[comp][off] = Sv * inv;
          if (a.outT[comp]) a.outT[comp][off] = Tv * inv;
        }
      }
    }
  }
}

extern "C" int mh_leg_analysis(int vector, const MhAnaArgs* args, void* stream) {
  const AnaArgs& a = *args;
  if (a.ncomp < 1 || a.ncomp > 3 || a.B < 1 || a.ntiles < 1 || a.Hp % 32 != 0 || a.Hp < a.H ||
      a.n_m < 1)
    return MH_EINVAL;
  const int n_m = a.n_m;
  dim3 grid(a.ncomp * mh_cdiv(a.B, 16), a.ntiles);
  if (a.prec != 0 && a.prec != 1) return MH_EINVAL;
  hipStream_t st = mh_stream(stream);
  if (a.prec == 0) {
    if (vector) hipLaunchKernelGGL((k_ana<true, double>), grid, dim3(256), 0, st, a, n_m);
    else hipLaunchKernelGGL((k_ana<false, double>), grid, dim3(256), 0, st, a, n_m);
  } else {
    if (vector) hipLaunchKernelGGL((k_ana<true, float>), grid, dim3(256), 0, st, a, n_m);
    else hipLaunchKernelGGL((k_ana<false, float>), grid, dim3(256), 0, st, a, n_m);
  }
  return mh_launch_status();
}


// ---------------------------------------------------------------- FP32 tables
__global__ void k_to_f32(const double* __restrict__ src, float* __restrict__ dst, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

extern "C" int mh_tables_to_f32(const double* src, float* dst, long n, void* stream) {
  if (n < 0) return MH_EINVAL;
  if (n == 0) return MH_OK;
  const long nb = (n + 255) / 256;
  if (nb > 0x7fffffffL) return MH_EINVAL;
  hipLaunchKernelGGL(k_to_f32, dim3((unsigned)nb), dim3(256), 0, mh_stream(stream), src, dst, n);
  return mh_launch_status();
}
