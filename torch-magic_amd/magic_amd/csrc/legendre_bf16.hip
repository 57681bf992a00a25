// Split-precision Legendre synthesis and analysis on the BF16 matrix cores
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation) for gfx950.
//
// Every operand x (table value or staged input) is held as NP bf16 planes,
// x ~ x_0 + x_1 + ... + x_{NP-1}, x_0 = bf16(x) and x_i = bf16 of the
// remainder, i.e. about 8 NP significant bits.  A product a*b is formed from
// the plane products a_i b_j with i + j < NP (NP = 2: hi*hi + hi*lo + lo*hi,
// three MFMAs; NP = 3: six MFMAs), the dropped terms being below
// 2^-(8 NP) |a b|; the MFMAs accumulate in FP32.  One BF16 MFMA contracts 32
// degrees (synthesis) or 32 theta pairs (analysis), i.e. a whole LDS chunk,
// where the FP32 kernels need eight 16x16x4 steps.
//
// Operand layouts of the 16x16x32 BF16 MFMA: A (16 x 32): lane l holds row
// l % 16, contraction indices 8 (l / 16) .. 8 (l / 16) + 7 (16 bytes); B
// (32 x 16): lane l holds column l % 16, the same contraction indices; C/D
// (16 x 16, FP32): element v of lane l is row 4 (l / 16) + v, column l % 16
// (the layout of the FP32 16x16x4 MFMA).  Tables come from mh_tables_to_bf16
// in layouts that make each lane's A operand 16 contiguous bytes; the staged
// inputs are written to LDS as [plane][column][32 (+8 padding)] bf16 rows so
// that each lane's B operand is one 16-byte LDS read.
#include "legendre_common.h"

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));

namespace {

// IEEE float -> bf16 bits, round to nearest even (finite inputs).
__device__ __forceinline__ unsigned bf16_rn(float x) {
  const unsigned u = __float_as_uint(x);
  return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf16_val(unsigned h) { return __uint_as_float(h << 16); }

template <int NP>
__device__ __forceinline__ void split_bf16(float x, unsigned short (&h)[NP]) {
  float r = x;
#pragma unroll
  for (int i = 0; i < NP; ++i) {
    const unsigned b = bf16_rn(r);
    h[i] = (unsigned short)b;
    r -= bf16_val(b);  // exact: the remainder has at most 16 significant bits
  }
}

__device__ __forceinline__ float4_t mfma_bf16(uint4 a, uint4 b, float4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                 __builtin_bit_cast(bf16x8_t, b), c, 0, 0, 0);
}

// sum of the plane products a_i b_j, i + j < NP, smallest terms first
template <int NP>
__device__ __forceinline__ float4_t mfma_split(const uint4 (&A)[NP], const uint4 (&B)[NP],
                                               float4_t c) {
#pragma unroll
  for (int s = NP - 1; s >= 0; --s)
#pragma unroll
    for (int i = 0; i <= s; ++i) c = mfma_bf16(A[i], B[s - i], c);
  return c;
}

constexpr int kRow16 = 40;  // LDS row: 32 bf16 values + 8 padding (80 bytes)

}  // namespace

// ---------------------------------------------------------------- synthesis
// Same work split as k_syn: one workgroup per (order m, 64 theta pairs,
// 8 NCT complex columns), wave w takes theta pairs j0 + 16 w; per parity and
// per chunk of 32 degrees the inputs (with the l(l+1) / or2 prefactors of
// step_time.py:986-1015) are split and staged in LDS, double-buffered, and
// the table planes of the next chunk are loaded while the current chunk's
// MFMAs run.
template <bool VEC, int NP, int NCT, typename FT>
__global__ __launch_bounds__(256, 2) void k_syn_bf(SynArgs a, int n_m) {
  constexpr int WAVES = 4;
  FT* const Fo = reinterpret_cast<FT*>(a.F);
  constexpr int NF = VEC ? 3 : 1;
  constexpr int NT = 64 * WAVES;
  constexpr int kSynJ = 16 * WAVES;
  constexpr int CW = 16 * NCT;
  constexpr int RPT = NT / CW;
  constexpr int NPS = (kChunk + RPT - 1) / RPT;
  constexpr int NA = (VEC ? 10 : 2) * NCT;
  constexpr int NTAB = VEC ? 2 : 1;
  static_assert(NT % CW == 0, "staging layout");
  __shared__ __attribute__((aligned(16))) unsigned short xs[2][NF][NP][CW][kRow16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nx = (a.ngroups * a.B + 8 * NCT - 1) / (8 * NCT), ny = (a.H + kSynJ - 1) / kSynJ;
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int m = (slot / (nx * ny)) * 8 + xcd;
  if (m >= n_m) return;  // uniform per workgroup
  const int mo = m * (a.minc > 1 ? a.minc : 1);  // the order of mode index m
  const int rem = slot % (nx * ny), ty = rem / nx, tx = rem - ty * nx;
  const int bc0 = tx * 8 * NCT;
  const int j0 = ty * kSynJ + 16 * w;
  const int col = lane & 15, kq = lane >> 4;
  const long lms = a.lm_start[m];
  const int ncol = a.ngroups * a.B;
  const int cw = threadIdx.x % CW, srow = threadIdx.x / CW;
  const int sbc = bc0 + (cw >> 1);
  const bool scol = sbc < ncol;
  const int sg = scol ? sbc / a.B : 0;
  const int sb = scol ? sbc - sg * a.B : 0;
  const long soff = 2L * sb + (cw & 1);
  const double sor2 = scol ? a.or2[sb] : 0.0;
  const double* sx1[NF];
  const double* sx2 = nullptr;
  int smodes = 0;
#pragma unroll
  for (int f = 0; f < NF; ++f) {
    const SynIn in = syn_in(a, sg, f);
    sx1[f] = in.x1 + soff;
    if (f == 2) sx2 = in.x2 + soff;
    smodes |= in.mode << (2 * f);
  }
  float4_t acc[NA];
#pragma unroll
  for (int q = 0; q < NA; ++q) acc[q] = (float4_t){0, 0, 0, 0};
  double r1[NF][NPS], r2[NPS];
  const bool wact = j0 < a.H;  // uniform per wave; then j0 + 15 < Hp
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const int* info = a.mp_info + (m * 2 + p) * 4;
    const int K = info[0];
    if (K <= 0) continue;
    const int K32 = a.mp16[(m * 2 + p) * 4];
    const long so = a.mp16[(m * 2 + p) * 4 + 1];
    const long ps = (long)a.Hp * K32;  // plane stride
    const unsigned short* tb[NTAB];
    tb[0] = reinterpret_cast<const unsigned short*>(a.synP) + so + (long)(j0 + col) * K32 + 8 * kq;
    if (VEC) tb[NTAB - 1] = reinterpret_cast<const unsigned short*>(a.synD) + so +
                            (long)(j0 + col) * K32 + 8 * kq;
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
            unsigned short h[NP];
            split_bf16<NP>((float)x, h);
#pragma unroll
            for (int pl = 0; pl < NP; ++pl) xs[buf][f][pl][cw][kr] = h[pl];
          }
        }
      }
    };
    fetch(0);
    __syncthreads();  // previous parity's readers are done with both buffers
    store(0, 0);
    __syncthreads();
    const int nch = (K + kChunk - 1) / kChunk;
    // table planes of the current and the next chunk
    uint4 tcur[NTAB][NP], tnxt[NTAB][NP];
    auto tload = [&](uint4 (&ta)[NTAB][NP], int c) {
#pragma unroll
      for (int t = 0; t < NTAB; ++t)
#pragma unroll
        for (int pl = 0; pl < NP; ++pl)
          ta[t][pl] = *reinterpret_cast<const uint4*>(tb[t] + pl * ps + 32L * c);
    };
    if (wact) tload(tcur, 0);
    for (int ch = 0; ch < nch; ++ch) {
      const int buf = ch & 1;
      const bool more = ch + 1 < nch;
      if (more) fetch((ch + 1) * kChunk);
      if (wact) {
        if (more) tload(tnxt, ch + 1);
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) {
          const int cc = 16 * ct + col;
          uint4 xq[NP], xsv[NP], xt[NP];
#pragma unroll
          for (int pl = 0; pl < NP; ++pl) {
            xq[pl] = *reinterpret_cast<const uint4*>(&xs[buf][0][pl][cc][8 * kq]);
            if (VEC) {
              xsv[pl] = *reinterpret_cast<const uint4*>(&xs[buf][1][pl][cc][8 * kq]);
              xt[pl] = *reinterpret_cast<const uint4*>(&xs[buf][NF - 1][pl][cc][8 * kq]);
            }
          }
          const uint4(&aP)[NP] = tcur[0];
          if (!VEC) {
            acc[p * NCT + ct] = mfma_split<NP>(aP, xq, acc[p * NCT + ct]);
          } else {
            const uint4(&aD)[NP] = tcur[NTAB - 1];
            acc[(5 * p + 0) * NCT + ct] = mfma_split<NP>(aP, xq, acc[(5 * p + 0) * NCT + ct]);   // P Q
            acc[(5 * p + 1) * NCT + ct] = mfma_split<NP>(aP, xt, acc[(5 * p + 1) * NCT + ct]);   // P T
            acc[(5 * p + 2) * NCT + ct] = mfma_split<NP>(aP, xsv, acc[(5 * p + 2) * NCT + ct]);  // P S
            acc[(5 * p + 3) * NCT + ct] = mfma_split<NP>(aD, xsv, acc[(5 * p + 3) * NCT + ct]);  // dP S
            acc[(5 * p + 4) * NCT + ct] = mfma_split<NP>(aD, xt, acc[(5 * p + 4) * NCT + ct]);   // dP T
          }
        }
      }
      if (more) {
        store(buf ^ 1, (ch + 1) * kChunk);
        __syncthreads();
#pragma unroll
        for (int t = 0; t < NTAB; ++t)
#pragma unroll
          for (int pl = 0; pl < NP; ++pl) tcur[t][pl] = tnxt[t][pl];
      }
    }
  }
  if (!wact) return;
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
      const int j = j0 + mfma_row_f32(lane, i);
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

int leg_synth_bf16(int vector, const SynArgs& a, int n_m, hipStream_t st) {
  const int nct = a.nct == 0 ? 1 : a.nct;
  if (a.mp16 == nullptr || (a.waves != 0 && a.waves != 4) || (nct == 2 && !vector) ||
      (nct != 1 && nct != 2))
    return MH_EINVAL;
  const long units = (long)mh_cdiv(a.ngroups * a.B, 8 * nct) * mh_cdiv(a.H, 64) * mh_cdiv(n_m, 8) * 8;
  if (units > 0x7fffffffL) return MH_EINVAL;
  const dim3 grid((unsigned)units), blk(256);
#define MH_SYNBF(VEC, NP, NCT)                                                               \
  if (a.fprec) hipLaunchKernelGGL((k_syn_bf<VEC, NP, NCT, float>), grid, blk, 0, st, a, n_m); \
  else hipLaunchKernelGGL((k_syn_bf<VEC, NP, NCT, double>), grid, blk, 0, st, a, n_m);
  if (a.prec == 2) {
    if (!vector) { MH_SYNBF(false, 2, 1) }
    else if (nct == 1) { MH_SYNBF(true, 2, 1) }
    else { MH_SYNBF(true, 2, 2) }
  } else {
    if (!vector) { MH_SYNBF(false, 3, 1) }
    else if (nct == 1) { MH_SYNBF(true, 3, 1) }
    else { MH_SYNBF(true, 3, 2) }
  }
#undef MH_SYNBF
  return mh_launch_status();
}

// ---------------------------------------------------------------- analysis
// Same work split as k_ana with four waves: one workgroup per (tile (m, p, k0)
// of 64 degrees, component, 8 complex columns), wave w takes the 16 degrees
// k0 + 16 w; per chunk of 32 theta pairs the hemisphere folds are split and
// staged in LDS (double-buffered) and the weight-table planes of the next
// chunk are loaded during the current chunk's MFMAs.
template <bool VEC, int NP, typename FT>
__global__ __launch_bounds__(256, 2) void k_ana_bf(AnaArgs a, int n_m) {
  constexpr int WAVES = 4;
  const FT* const Fi = reinterpret_cast<const FT*>(a.F);
  constexpr int NF = VEC ? 4 : 1;
  constexpr int NT = 64 * WAVES;
  constexpr int NTAB = VEC ? 2 : 1;
  __shared__ __attribute__((aligned(16))) unsigned short fs[2][NF][NP][16][kRow16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
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
  const int col = lane & 15, kq = lane >> 4;
  const double sg = p ? -1.0 : 1.0;
  const int* info = a.mp_info + (m * 2 + p) * 4;
  const int K = info[0], K16 = info[3];
  const long so = a.mp16[(m * 2 + p) * 4 + 2];
  const long ps = (long)K16 * a.Hp;  // plane stride
  const int c1 = a.cf1[comp], c2 = VEC ? a.cf2[comp] : 0;
  float4_t acc[VEC ? 4 : 1];
#pragma unroll
  for (int q = 0; q < (VEC ? 4 : 1); ++q) acc[q] = (float4_t){0, 0, 0, 0};
  constexpr int RP = NT / 16;
  static_assert(2 * RP >= kChunk && RP <= kChunk, "two staging passes cover a 32-row chunk");
  const int c16 = threadIdx.x & 15, srow = threadIdx.x >> 4;
  const bool s2 = RP + srow < kChunk;
  const int sb = bc0 + (c16 >> 1);
  const bool scol = sb < a.B;
  const FT* f1p = Fi + c1 * a.c_stride + fidx(0, a.m_stride, a.Bin, 0, m, 0, scol ? sb : 0) + (c16 & 1);
  const FT* f2p = Fi + c2 * a.c_stride + fidx(0, a.m_stride, a.Bin, 0, m, 0, scol ? sb : 0) + (c16 & 1);
  const long rstride = 2L * a.Bin;
  double raw[2][VEC ? 4 : 2];
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
  auto put = [&](int buf, int f, int jr, double v) {
    unsigned short h[NP];
    split_bf16<NP>((float)v, h);
#pragma unroll
    for (int pl = 0; pl < NP; ++pl) fs[buf][f][pl][c16][jr] = h[pl];
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      if (q == 0 || s2) {
        const int jr = RP * q + srow;
        put(buf, 0, jr, raw[q][0] + sg * raw[q][1]);
        if (VEC) {
          put(buf, 1, jr, raw[q][0] - sg * raw[q][1]);
          put(buf, 2, jr, raw[q][2] + sg * raw[q][3]);
          put(buf, 3, jr, raw[q][2] - sg * raw[q][3]);
        }
      }
    }
  };
  const bool active = kbase < K;  // uniform per wave; then kbase + 15 < K16
  const unsigned short* tb[NTAB];
  tb[0] = reinterpret_cast<const unsigned short*>(a.anaW) + so + (long)(kbase + col) * a.Hp + 8 * kq;
  if (VEC) tb[NTAB - 1] = reinterpret_cast<const unsigned short*>(a.anaWd) + so +
                          (long)(kbase + col) * a.Hp + 8 * kq;
  uint4 tcur[NTAB][NP], tnxt[NTAB][NP];
  auto tload = [&](uint4 (&ta)[NTAB][NP], int c) {
#pragma unroll
    for (int t = 0; t < NTAB; ++t)
#pragma unroll
      for (int pl = 0; pl < NP; ++pl)
        ta[t][pl] = *reinterpret_cast<const uint4*>(tb[t] + pl * ps + 32L * c);
  };
  fetch(0);
  store(0);
  __syncthreads();
  const int nch = (a.H + kChunk - 1) / kChunk;  // Hp >= 32 nch: table rows exist
  if (active) tload(tcur, 0);
  for (int ch = 0; ch < nch; ++ch) {
    const int buf = ch & 1;
    const bool more = ch + 1 < nch;
    if (more) fetch((ch + 1) * kChunk);
    if (active) {
      if (more) tload(tnxt, ch + 1);
      uint4 x1[NP], x2[NP], x3[NP], x4[NP];
#pragma unroll
      for (int pl = 0; pl < NP; ++pl) {
        x1[pl] = *reinterpret_cast<const uint4*>(&fs[buf][0][pl][col][8 * kq]);
        if (VEC) {
          x2[pl] = *reinterpret_cast<const uint4*>(&fs[buf][1][pl][col][8 * kq]);
          x3[pl] = *reinterpret_cast<const uint4*>(&fs[buf][2][pl][col][8 * kq]);
          x4[pl] = *reinterpret_cast<const uint4*>(&fs[buf][3][pl][col][8 * kq]);
        }
      }
      const uint4(&aw)[NP] = tcur[0];
      if (!VEC) {
        acc[0] = mfma_split<NP>(aw, x1, acc[0]);
      } else {
        const uint4(&ad)[NP] = tcur[NTAB - 1];
        acc[0] = mfma_split<NP>(ad, x4, acc[0]);  // wdP f2 (opposite fold)
        acc[1] = mfma_split<NP>(ad, x2, acc[1]);  // wdP f1 (opposite fold)
        acc[2] = mfma_split<NP>(aw, x1, acc[2]);  // wP f1 (same fold)
        acc[3] = mfma_split<NP>(aw, x3, acc[3]);  // wP f2 (same fold)
      }
    }
    if (more) {
      store(buf ^ 1);
      __syncthreads();
#pragma unroll
      for (int t = 0; t < NTAB; ++t)
#pragma unroll
        for (int pl = 0; pl < NP; ++pl) tcur[t][pl] = tnxt[t][pl];
    }
  }
  if (!active) return;
  const double dm = (double)mo;
  const long lms = a.lm_start[m];
  const int b = bc0 + (col >> 1), reim = col & 1;
  const bool bval = b < a.B;
  const int bl = bval ? b : 0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int k = kbase + mfma_row_f32(lane, i);
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

int leg_analysis_bf16(int vector, const AnaArgs& a, hipStream_t st) {
  if (a.mp16 == nullptr || a.waves != 4) return MH_EINVAL;
  const dim3 grid((unsigned)(a.ncomp * mh_cdiv(a.B, 8) * a.ntiles)), blk(256);
#define MH_ANABF(VEC, NP)                                                                    \
  if (a.fprec) hipLaunchKernelGGL((k_ana_bf<VEC, NP, float>), grid, blk, 0, st, a, a.n_m);   \
  else hipLaunchKernelGGL((k_ana_bf<VEC, NP, double>), grid, blk, 0, st, a, a.n_m);
  if (a.prec == 2) {
    if (vector) { MH_ANABF(true, 2) } else { MH_ANABF(false, 2) }
  } else {
    if (vector) { MH_ANABF(true, 3) } else { MH_ANABF(false, 3) }
  }
#undef MH_ANABF
  return mh_launch_status();
}

// ---------------------------------------------------------------- BF16 tables
// One workgroup row per (m, parity) block; kind 0: synthesis table, 1: analysis.
__global__ void k_tables_bf16(const double* __restrict__ src, unsigned short* __restrict__ dst,
                              const int* __restrict__ mp_info, const int* __restrict__ mp16,
                              int H, int Hp, int np, int kind) {
  const int mp = blockIdx.y;
  const int K = mp_info[mp * 4], K16 = mp_info[mp * 4 + 3];
  if (K <= 0) return;
  const int K32 = mp16[mp * 4];
  const long n = kind == 0 ? (long)Hp * K32 : (long)K16 * Hp;
  const long so = kind == 0 ? mp_info[mp * 4 + 1] : mp_info[mp * 4 + 2];
  const long dof = kind == 0 ? mp16[mp * 4 + 1] : mp16[mp * 4 + 2];
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    double x = 0.0;
    if (kind == 0) {  // dst [j][k], src [k][j] (k < K4)
      const int j = (int)(e / K32), k = (int)(e - (long)j * K32);
      if (k < K) x = src[so + (long)k * Hp + j];
    } else {  // dst [k][j], src [j][k] (k < K16)
      const int k = (int)(e / Hp), j = (int)(e - (long)k * Hp);
      if (k < K && j < H) x = src[so + (long)j * K16 + k];
    }
    double r = x;
    for (int pl = 0; pl < np; ++pl) {
      const unsigned b = bf16_rn((float)r);
      dst[dof + pl * n + e] = (unsigned short)b;
      r -= (double)bf16_val(b);
    }
  }
}

extern "C" int mh_tables_to_bf16(const double* synP, const double* synD, const double* anaW,
                                 const double* anaWd, const int* mp_info, const int* mp16,
                                 int n_mp, int H, int Hp, int np, unsigned short* synP16,
                                 unsigned short* synD16, unsigned short* anaW16,
                                 unsigned short* anaWd16, void* stream) {
  if (n_mp < 1 || H < 1 || Hp < H || Hp % 32 != 0 || (np != 2 && np != 3) || !mp_info || !mp16)
    return MH_EINVAL;
  hipStream_t st = mh_stream(stream);
  const dim3 grid(64, (unsigned)n_mp), blk(256);
  const double* srcs[4] = {synP, synD, anaW, anaWd};
  unsigned short* dsts[4] = {synP16, synD16, anaW16, anaWd16};
  for (int t = 0; t < 4; ++t) {
    if (srcs[t] == nullptr || dsts[t] == nullptr) continue;
    hipLaunchKernelGGL(k_tables_bf16, grid, blk, 0, st, srcs[t], dsts[t], mp_info, mp16, H, Hp,
                       np, t < 2 ? 0 : 1);
  }
  return mh_launch_status();
}
