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
#include "legendre_common.h"

// ---------------------------------------------------------------- P_lm tables
// Orthonormal P_lm without Condon-Shortley phase and sin(theta) dP_lm/dtheta by
// the three-term recurrence (the same recurrences as the reference plms.py).
__global__ void k_plm_tables(int l_max, int minc, int H, int Hp, const double* __restrict__ cos_t,
                             const double* __restrict__ sin_t, const double* __restrict__ wgt,
                             const int* __restrict__ mp_info, double* __restrict__ synP,
                             double* __restrict__ synD, double* __restrict__ anaW,
                             double* __restrict__ anaWd) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int mi = blockIdx.y, m = mi * minc;  // mode index, order
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
    const int* info = mp_info + (mi * 2 + p) * 4;
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

extern "C" int mh_plm_tables(int l_max, int n_m, int minc, int H, int Hp, const double* cos_t,
                             const double* sin_t, const double* wgt, const int* mp_info,
                             double* synP, double* synD, double* anaW, double* anaWd,
                             void* stream) {
  if (l_max < 0 || minc < 1 || n_m != l_max / minc + 1 || H <= 0 || Hp < H) return MH_EINVAL;
  dim3 grid(mh_cdiv(H, 64), n_m);
  hipLaunchKernelGGL(k_plm_tables, grid, dim3(64), 0, mh_stream(stream), l_max, minc, H, Hp, cos_t,
                     sin_t, wgt, mp_info, synP, synD, anaW, anaWd);
  return mh_launch_status();
}

// ---------------------------------------------------------------- synthesis

__device__ __forceinline__ double syn_load(const SynIn& in, long lm, int l, int b, int reim,
                                           int ldx, const double* or2) {
  const long off = lm * 2 * (long)ldx + 2 * b + reim;
  const double dl = (double)l * (double)(l + 1);
  if (in.mode == 0) return in.x1[off];
  if (in.mode == 1) return dl * in.x1[off];
  return -in.x1[off] + or2[b] * dl * in.x2[off];
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

// Synthesis, one workgroup per (order m, 96 theta pairs, 8 complex columns).
// Columns of the groups are laid side by side (column bc is batch column
// bc % B of group bc / B, which is also its column in F).  The input tile of
// 32 degrees x 16 real columns x (1 or 3) fields is staged in LDS with the
// l(l+1) / or2 prefactors applied once per element, double-buffered: the loads
// of the next chunk are issued before the current chunk's MFMAs and combined
// and stored after them.  Wave w takes theta pairs j0 + 16w and streams its
// table operands from global memory through a ring of kSynDepth steps.
// Waves per synthesis workgroup: each wave takes 16 theta pairs.  Four (the
// default) or six; both give identical results, the tuner may pick either.
constexpr int kSynWavesDefault = 4, kSynDepth = 4;

// NCT column tiles of 8 complex columns per workgroup: every table operand
// feeds NCT MFMAs.  Two are instantiated for the FP32 vector synthesis (its
// 20 accumulators fit; FP64 would not), chosen by the tuner.
template <bool VEC, typename T, int NCT, typename FT, int WAVES>
__global__ __launch_bounds__(64 * WAVES, 3) void k_syn(SynArgs a, int n_m) {
  FT* const Fo = reinterpret_cast<FT*>(a.F);  // Fourier buffer elements (double or float)
  typedef Mfma<T> MF;
  typedef typename MF::acc_t acc_t;
  constexpr int NF = VEC ? 3 : 1;           // fields per group (Q, S, T)
  constexpr int NT = 64 * WAVES;            // threads
  constexpr int kSynJ = 16 * WAVES;         // theta pairs per workgroup
  constexpr int SPC = kChunk / 4;           // MFMA steps per chunk
  constexpr int D = kSynDepth;
  constexpr int CW = 16 * NCT;              // real columns per workgroup
  constexpr int RPT = NT / CW;              // chunk rows staged per pass
  constexpr int NPS = (kChunk + RPT - 1) / RPT;  // staging passes
  constexpr int NA = (VEC ? 10 : 2) * NCT;  // accumulators
  static_assert(NT % CW == 0 && SPC % D == 0, "staging / ring layout");
  __shared__ T xs[2][NF][kChunk][CW];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // XCD-aware order: workgroups b and b+8 run on the same XCD; XCD x takes the
  // orders m = x, x+8, ... (all tiles of one m on one XCD, so the m's tables
  // are fetched into that XCD's L2 once), which also balances the K ~ l_max-m work
  const int nx = (a.ngroups * a.B + 8 * NCT - 1) / (8 * NCT), ny = (a.H + kSynJ - 1) / kSynJ;
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int m = (slot / (nx * ny)) * 8 + xcd;
  if (m >= n_m) return;  // uniform per workgroup
  const int mo = m * (a.minc > 1 ? a.minc : 1);  // the order of mode index m
  const int rem = slot % (nx * ny), ty = rem / nx, tx = rem - ty * nx;
  const int bc0 = tx * 8 * NCT;
  const int j0 = ty * kSynJ + 16 * w;
  const int col = lane & 15, kk = lane >> 4;
  const long lms = a.lm_start[m];
  const int ncol = a.ngroups * a.B;
  // staging map: thread stages real column cw = tid % CW of chunk rows
  // tid / CW + RPT*q (q < NPS, rows < kChunk) of every field
  const int cw = threadIdx.x % CW, srow = threadIdx.x / CW;
  const int sbc = bc0 + (cw >> 1);
  const bool scol = sbc < ncol;
  const int sg = scol ? sbc / a.B : 0;
  const int sb = scol ? sbc - sg * a.B : 0;
  const long soff = 2L * sb + (cw & 1);
  const double sor2 = scol ? a.or2[sb] : 0.0;
  // this thread's input pointers (column offset folded in) and modes, resolved
  // once: per-lane selection inside the chunk loop would re-read the argument
  // block through memory on every chunk
  const double* sx1[NF];
  const double* sx2 = nullptr;
  int smodes = 0;  // 2 bits per field
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const SynIn in = syn_in(a, sg, f);
    sx1[f] = in.x1 + soff;
    if (f == 2) sx2 = in.x2 + soff;
    smodes |= in.mode << (2 * f);
  }
  acc_t acc[NA];  // acc[(5p + q) * NCT + ct] (vector), acc[p * NCT + ct] (scalar)
#pragma unroll
  for (int q = 0; q < NA; ++q) acc[q] = (acc_t){0, 0, 0, 0};
  double r1[NF][NPS], r2[NPS];  // raw inputs of the next chunk (r2: second input of T)
  const bool wact = j0 < a.H;  // uniform per wave; then j0 + 15 < Hp
#pragma unroll  // static accumulator indices (acc[] stays in registers)
  for (int p = 0; p < 2; ++p) {
    const int* info = a.mp_info + (m * 2 + p) * 4;
    const int K = info[0];
    if (K <= 0) continue;
    const T* tP = reinterpret_cast<const T*>(a.synP) + info[1];
    const T* tD = reinterpret_cast<const T*>(a.synD) + info[1];
    auto fetch = [&](int c0) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
#pragma unroll
        for (int q = 0; q < NPS; ++q) {
          const int kr = srow + RPT * q, k = c0 + kr;
          r1[f][q] = 0.0;
          if (f == 2) r2[q] = 0.0;
          if (kr < kChunk && scol && k < K) {
            const long o = (lms + p + 2 * k) * 2 * (long)a.ldx;
            r1[f][q] = sx1[f][o];
            if (f == 2 && ((smodes >> 4) & 3) == 2) r2[q] = sx2[o];
          }
        }
      }
    };
    auto store = [&](int buf, int c0) {
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const int mode = (smodes >> (2 * f)) & 3;
#pragma unroll
        for (int q = 0; q < NPS; ++q) {
          const int kr = srow + RPT * q;
          if (kr < kChunk) {
            const int l = mo + p + 2 * (c0 + kr);
            const double dl = (double)l * (double)(l + 1);
            const double x1 = r1[f][q], x2 = (f == 2) ? r2[q] : 0.0;
            const double x = mode == 0 ? x1 : mode == 1 ? dl * x1 : -x1 + sor2 * dl * x2;
            xs[buf][f][kr][cw] = (T)x;
          }
        }
      }
    };
    fetch(0);
    __syncthreads();  // previous parity's readers are done with both buffers
    store(0, 0);
    __syncthreads();
    // table rows are padded to a multiple of 4; waves past the last theta pair
    // only help with the staging
    const int nsteps = wact ? (K + 3) / 4 : 0;
    const T* pP = tP + (long)kk * a.Hp + j0 + col;
    const T* pD = tD + (long)kk * a.Hp + j0 + col;
    const long tstep = 4L * a.Hp;
    T rP[D], rD[D];  // table operands of steps t .. t+D-1 (slot t % D)
#pragma unroll
    for (int d = 0; d < D; ++d) {
      rP[d] = (T)0;
      rD[d] = (T)0;
      if (d < nsteps) {
        rP[d] = pP[d * tstep];
        if (VEC) rD[d] = pD[d * tstep];
      }
    }
    const int nch = (K + kChunk - 1) / kChunk;
    for (int ch = 0; ch < nch; ++ch) {
      const int buf = ch & 1;
      const bool more = ch + 1 < nch;
      // one MFMA step: degrees 4t .. 4t+3 against LDS rows 4s .. 4s+3; the
      // slot is refilled with step t + D after its last use, straight from
      // global memory into the ring register (row clamped to the table)
      auto step = [&](const int s, const int t) {
        const T aP = rP[s % D], aD = rD[s % D];
        const int kr = 4 * s + kk;
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) {
          const int cc = 16 * ct + col;
          const T xq = xs[buf][0][kr][cc];
          if (!VEC) {
            acc[p * NCT + ct] = MF::run(aP, xq, acc[p * NCT + ct]);
          } else {
            const T xsv = xs[buf][1][kr][cc];
            const T xt = xs[buf][2][kr][cc];
            acc[(5 * p + 0) * NCT + ct] = MF::run(aP, xq, acc[(5 * p + 0) * NCT + ct]);   // P Q
            acc[(5 * p + 1) * NCT + ct] = MF::run(aP, xt, acc[(5 * p + 1) * NCT + ct]);   // P T
            acc[(5 * p + 2) * NCT + ct] = MF::run(aP, xsv, acc[(5 * p + 2) * NCT + ct]);  // P S
            acc[(5 * p + 3) * NCT + ct] = MF::run(aD, xsv, acc[(5 * p + 3) * NCT + ct]);  // dP S
            acc[(5 * p + 4) * NCT + ct] = MF::run(aD, xt, acc[(5 * p + 4) * NCT + ct]);   // dP T
          }
        }
        const long tn = (long)min(t + D, nsteps - 1) * tstep;
        rP[s % D] = pP[tn];
        if (VEC) rD[s % D] = pD[tn];
      };
      // the next chunk's inputs are requested after the first step, so that
      // step's operands are not queued behind them
      // (the staging store is repeated in both branches so that its wait
      // covers only the staged loads, not the table refills issued after them)
      if ((ch + 1) * SPC <= nsteps) {  // full chunk (uniform per wave): straight-line steps
        step(0, ch * SPC);
        if (more) fetch((ch + 1) * kChunk);
#pragma unroll
        for (int s = 1; s < SPC; ++s) step(s, ch * SPC + s);
        if (more) store(buf ^ 1, (ch + 1) * kChunk);
      } else {  // last chunk of an active wave, or an idle wave
        if (more) fetch((ch + 1) * kChunk);
#pragma unroll
        for (int s = 0; s < SPC; ++s)
          if (ch * SPC + s < nsteps) step(s, ch * SPC + s);
        if (more) store(buf ^ 1, (ch + 1) * kChunk);
      }
      if (more) __syncthreads();
    }
  }
  const double dm = (double)mo;
  const int reim = col & 1;
#pragma unroll
  for (int ct = 0; ct < NCT; ++ct) {
  const int bc = bc0 + 8 * ct + (col >> 1);
  const bool bval = bc < ncol;
  const int bo = bval ? bc : 0;
  auto A = [&](int q, int i) { return (double)acc[q * NCT + ct][i]; };
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int j = j0 + MF::row(lane, i);
    const bool st = bval && (j < a.H);
    const long oN = fidx(a.c_stride, a.m_stride, a.Bout, 0, m, 2 * j, bo) + reim;
    const long oS = fidx(a.c_stride, a.m_stride, a.Bout, 0, m, 2 * j + 1, bo) + reim;
    if (!VEC) {
      const double e = A(0, i), o = A(1, i);
      if (st) {
        Fo[oN] = (FT)(e + o);
        Fo[oS] = (FT)(e - o);
      }
    } else {
      const double Qe = A(0, i), Qo = A(5, i);
      const double PTe = A(1, i), PTo = A(6, i);
      const double PSe = A(2, i), PSo = A(7, i);
      const double DSe = A(3, i), DSo = A(8, i);
      const double DTe = A(4, i), DTo = A(9, i);
      const double iPTp = imul_pair(PTe + PTo, dm, lane), iPTm = imul_pair(PTe - PTo, dm, lane);
      const double iPSp = imul_pair(PSe + PSo, dm, lane), iPSm = imul_pair(PSe - PSo, dm, lane);
      if (st) {
        Fo[oN] = (FT)(Qe + Qo);
        Fo[oS] = (FT)(Qe - Qo);
        Fo[a.c_stride + oN] = (FT)((DSe + DSo) + iPTp);
        Fo[a.c_stride + oS] = (FT)(-(DSe - DSo) + iPTm);
        Fo[2 * a.c_stride + oN] = (FT)(iPSp - (DTe + DTo));
        Fo[2 * a.c_stride + oS] = (FT)(iPSm + (DTe - DTo));
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
  if (a.prec < 0 || a.prec > 3) return MH_EINVAL;
  if (a.fprec != 0 && (a.fprec != 1 || a.prec < 1)) return MH_EINVAL;
  // mode 2 (two inputs) is supported for the T field of a vector group only
  for (int g = 0; g < a.ngroups; ++g)
    for (int f = 0; f < (vector ? 3 : 1); ++f)
      if (a.in[g][f].mode < 0 || a.in[g][f].mode > 2 || (a.in[g][f].mode == 2 && !(vector && f == 2)))
        return MH_EINVAL;
  if (a.prec >= 2) return leg_synth_bf16(vector, a, n_m, mh_stream(stream));
  // column tiles per workgroup: one, or two for the FP32 vector synthesis
  const int nct = a.nct == 0 ? 1 : a.nct;
  if (nct != 1 && (nct != 2 || a.prec != 1 || !vector)) return MH_EINVAL;
  const int waves = a.waves == 0 ? kSynWavesDefault : a.waves;
  if (waves != 4 && waves != 6) return MH_EINVAL;
  const long units =
      (long)mh_cdiv(a.ngroups * a.B, 8 * nct) * mh_cdiv(a.H, 16 * waves) * mh_cdiv(n_m, 8) * 8;
  if (units > 0x7fffffffL) return MH_EINVAL;
  dim3 grid((unsigned)units);
  const dim3 blk(64 * waves);
  hipStream_t st = mh_stream(stream);
#define MH_SYN_LAUNCH(W)                                                                       \
  if (nct == 2) {                                                                              \
    if (a.fprec == 0) hipLaunchKernelGGL((k_syn<true, float, 2, double, W>), grid, blk, 0, st, a, n_m); \
    else hipLaunchKernelGGL((k_syn<true, float, 2, float, W>), grid, blk, 0, st, a, n_m);       \
  } else if (a.prec == 0) {                                                                    \
    if (vector) hipLaunchKernelGGL((k_syn<true, double, 1, double, W>), grid, blk, 0, st, a, n_m); \
    else hipLaunchKernelGGL((k_syn<false, double, 1, double, W>), grid, blk, 0, st, a, n_m);   \
  } else if (a.fprec == 0) {                                                                   \
    if (vector) hipLaunchKernelGGL((k_syn<true, float, 1, double, W>), grid, blk, 0, st, a, n_m); \
    else hipLaunchKernelGGL((k_syn<false, float, 1, double, W>), grid, blk, 0, st, a, n_m);    \
  } else {                                                                                     \
    if (vector) hipLaunchKernelGGL((k_syn<true, float, 1, float, W>), grid, blk, 0, st, a, n_m); \
    else hipLaunchKernelGGL((k_syn<false, float, 1, float, W>), grid, blk, 0, st, a, n_m);     \
  }
  if (waves == 4) {
    MH_SYN_LAUNCH(4)
  } else {
    MH_SYN_LAUNCH(6)
  }
#undef MH_SYN_LAUNCH
  return mh_launch_status();
}


// ---------------------------------------------------------------- analysis

// Analysis, one workgroup per (tile (m, p, k0), component, 8 complex columns):
// wave w takes the 16 degrees k0 + 16w (96 per tile).  The hemisphere folds
// of the Fourier rows (N + sg S and N - sg S, sg = +1 for even and -1 for odd
// l - m) for 32 theta pairs x 16 real columns are staged in LDS once per chunk
// (double-buffered) and shared by the six waves; the weight tables are read
// from global memory.
// Waves per analysis workgroup (16 degrees each): six (the default) or four;
// the tile list passed in must have tiles of 16 * waves degrees.
constexpr int kAnaWavesDefault = 6;

template <bool VEC, typename T, typename FT, int WAVES>
__global__ __launch_bounds__(64 * WAVES, 2) void k_ana(AnaArgs a, int n_m) {
  const FT* const Fi = reinterpret_cast<const FT*>(a.F);  // Fourier buffer elements
  typedef Mfma<T> MF;
  typedef typename MF::acc_t acc_t;
  constexpr int NF = VEC ? 4 : 1;          // f1 same, f1 opposite, f2 same, f2 opposite
  constexpr int NT = 64 * WAVES;
  __shared__ T fs[2][NF][kChunk][16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // XCD-aware order: the tile list holds 8 equal-length lists, list x with the
  // tiles of orders m = x (mod 8) (padding entries have m = n_m); workgroups
  // b and b+8 run on the same XCD, which works through its list in order
  const int nbt = (a.B + 7) / 8, gx = a.ncomp * nbt, tpx = a.ntiles / 8;
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int* tl = a.tiles + 3 * (xcd * tpx + slot / gx);
  const int m = tl[0], p = tl[1];
  if (m >= n_m) return;  // uniform per workgroup
  const int mo = m * (a.minc > 1 ? a.minc : 1);  // the order of mode index m
  const int kbase = tl[2] + 16 * w;
  const int cx = slot % gx;
  const int comp = cx / nbt;
  const int bc0 = (cx - comp * nbt) * 8;
  const int col = lane & 15, kk = lane >> 4;
  const double sg = p ? -1.0 : 1.0;
  const int* info = a.mp_info + (m * 2 + p) * 4;
  const int K = info[0], K16 = info[3];
  const T* tW = reinterpret_cast<const T*>(a.anaW) + info[2];
  const T* tD = reinterpret_cast<const T*>(a.anaWd) + info[2];
  const int c1 = a.cf1[comp], c2 = VEC ? a.cf2[comp] : 0;
  acc_t acc[VEC ? 4 : 1];
#pragma unroll
  for (int q = 0; q < (VEC ? 4 : 1); ++q) acc[q] = (acc_t){0, 0, 0, 0};
  // staging map: thread stages real column c16 = tid % 16 of theta pairs
  // RP*q + tid / 16 (q = 0, 1; RP = NT / 16 rows per pass; rows < 32) of the
  // chunk; the raw N/S values are loaded before the chunk's MFMAs and folded
  // after them
  constexpr int RP = NT / 16;
  static_assert(2 * RP >= kChunk && RP <= kChunk, "two staging passes cover a 32-row chunk");
  const int c16 = threadIdx.x & 15, srow = threadIdx.x >> 4;
  const bool s2 = RP + srow < kChunk;
  const int sb = bc0 + (c16 >> 1);
  const bool scol = sb < a.B;
  const FT* f1p = Fi + c1 * a.c_stride + fidx(0, a.m_stride, a.Bin, 0, m, 0, scol ? sb : 0) + (c16 & 1);
  const FT* f2p = Fi + c2 * a.c_stride + fidx(0, a.m_stride, a.Bin, 0, m, 0, scol ? sb : 0) + (c16 & 1);
  const long rstride = 2L * a.Bin;  // one grid row t
  double raw[2][VEC ? 4 : 2];       // (f1N, f1S, f2N, f2S) of the next chunk
  auto fetch = [&](int j0) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int j = j0 + RP * q + srow;
#pragma unroll
      for (int f = 0; f < (VEC ? 4 : 2); ++f) raw[q][f] = 0.0;
      if ((q == 0 || s2) && scol && j < a.H) {
        const long iN = 2L * j * rstride, iS = iN + rstride;
        raw[q][0] = f1p[iN];
        raw[q][1] = f1p[iS];
        if (VEC) {
          raw[q][2] = f2p[iN];
          raw[q][3] = f2p[iS];
        }
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      if (q == 0 || s2) {
        const int jr = RP * q + srow;
        fs[buf][0][jr][c16] = (T)(raw[q][0] + sg * raw[q][1]);
        if (VEC) {
          fs[buf][1][jr][c16] = (T)(raw[q][0] - sg * raw[q][1]);
          fs[buf][2][jr][c16] = (T)(raw[q][2] + sg * raw[q][3]);
          fs[buf][3][jr][c16] = (T)(raw[q][2] - sg * raw[q][3]);
        }
      }
    }
  };
  const bool active = kbase < K;  // uniform per wave; then kbase + 15 < K16
  // weight-table operands of theta step t (pairs 4t .. 4t+3) stream through a
  // ring of D steps; the tables hold Hp >= roundup(H, 4) rows
  constexpr int SPC = kChunk / 4, D = kSynDepth;
  const int nsteps = (a.H + 3) / 4;
  const T* pW = tW + (long)kk * K16 + kbase + col;
  const T* pD = tD + (long)kk * K16 + kbase + col;
  const long tstep = 4L * K16;
  T rW[D], rD[D];
#pragma unroll
  for (int d = 0; d < D; ++d) {
    rW[d] = (T)0;
    rD[d] = (T)0;
    if (active && d < nsteps) {
      rW[d] = pW[d * tstep];
      if (VEC) rD[d] = pD[d * tstep];
    }
  }
  fetch(0);
  store(0);
  __syncthreads();
  const int nch = (a.H + kChunk - 1) / kChunk;
  for (int ch = 0; ch < nch; ++ch) {
    const int buf = ch & 1;
    const bool more = ch + 1 < nch;
    auto step = [&](const int s, const int t) {
      const T aw = rW[s % D], ad = rD[s % D];
      const int jr = 4 * s + kk;
      const T x1 = fs[buf][0][jr][col];
      if (!VEC) {
        acc[0] = MF::run(aw, x1, acc[0]);
      } else {
        const T x2 = fs[buf][1][jr][col];
        const T x3 = fs[buf][2][jr][col];
        const T x4 = fs[buf][3][jr][col];
        acc[0] = MF::run(ad, x4, acc[0]);  // wdP f2 (opposite fold)
        acc[1] = MF::run(ad, x2, acc[1]);  // wdP f1 (opposite fold)
        acc[2] = MF::run(aw, x1, acc[2]);  // wP f1 (same fold)
        acc[3] = MF::run(aw, x3, acc[3]);  // wP f2 (same fold)
      }
      const long tn = (long)min(t + D, nsteps - 1) * tstep;
      rW[s % D] = pW[tn];
      if (VEC) rD[s % D] = pD[tn];
    };
    // buf ^ 1 was last read in chunk ch - 1, before the previous barrier; the
    // store is repeated per branch so its wait covers the staged loads only
    if (active && (ch + 1) * SPC <= nsteps) {  // full chunk: straight-line steps
      step(0, ch * SPC);
      if (more) fetch((ch + 1) * kChunk);
#pragma unroll
      for (int s = 1; s < SPC; ++s) step(s, ch * SPC + s);
      if (more) store(buf ^ 1);
    } else {
      if (more) fetch((ch + 1) * kChunk);
      if (active) {
#pragma unroll
        for (int s = 0; s < SPC; ++s)
          if (ch * SPC + s < nsteps) step(s, ch * SPC + s);
      }
      if (more) store(buf ^ 1);
    }
    if (more) __syncthreads();
  }
  if (!active) return;
  const double dm = (double)mo;
  const long lms = a.lm_start[m];
  const int b = bc0 + (col >> 1), reim = col & 1;
  const bool bval = b < a.B;
  const int bl = bval ? b : 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int k = kbase + MF::row(lane, i);
    const int l = mo + p + 2 * k;
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

extern "C" int mh_leg_analysis(int vector, const MhAnaArgs* args, void* stream) {
  const AnaArgs& a = *args;
  if (a.ncomp < 1 || a.ncomp > 3 || a.B < 1 || a.ntiles < 1 || a.Hp % 32 != 0 || a.Hp < a.H ||
      a.n_m < 1)
    return MH_EINVAL;
  const int n_m = a.n_m;
  if (a.ntiles % 8 != 0) return MH_EINVAL;
  const int waves = a.waves == 0 ? kAnaWavesDefault : a.waves;
  if (waves != 4 && waves != 6) return MH_EINVAL;
  dim3 grid((unsigned)(a.ncomp * mh_cdiv(a.B, 8) * a.ntiles));
  const dim3 blk(64 * waves);
  if (a.prec < 0 || a.prec > 3) return MH_EINVAL;
  if (a.fprec != 0 && (a.fprec != 1 || a.prec < 1)) return MH_EINVAL;
  hipStream_t st = mh_stream(stream);
  if (a.prec >= 2) return leg_analysis_bf16(vector, a, st);
#define MH_ANA_LAUNCH(W)                                                                      \
  if (a.prec == 0) {                                                                          \
    if (vector) hipLaunchKernelGGL((k_ana<true, double, double, W>), grid, blk, 0, st, a, n_m); \
    else hipLaunchKernelGGL((k_ana<false, double, double, W>), grid, blk, 0, st, a, n_m);     \
  } else if (a.fprec == 0) {                                                                  \
    if (vector) hipLaunchKernelGGL((k_ana<true, float, double, W>), grid, blk, 0, st, a, n_m); \
    else hipLaunchKernelGGL((k_ana<false, float, double, W>), grid, blk, 0, st, a, n_m);      \
  } else {                                                                                    \
    if (vector) hipLaunchKernelGGL((k_ana<true, float, float, W>), grid, blk, 0, st, a, n_m); \
    else hipLaunchKernelGGL((k_ana<false, float, float, W>), grid, blk, 0, st, a, n_m);       \
  }
  if (waves == 6) {
    MH_ANA_LAUNCH(6)
  } else {
    MH_ANA_LAUNCH(4)
  }
#undef MH_ANA_LAUNCH
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
