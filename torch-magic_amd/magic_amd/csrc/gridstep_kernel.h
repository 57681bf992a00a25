// Fused grid-space kernel of the radial loop (see gridstep.hip for the
// algorithm).  The kernel and its device helpers are shared by the
// translation units that instantiate it: gridstep.hip (each compiled plan's
// own level block, and the run-time plan), gridstep_rb1.hip and
// gridstep_rb2.hip (compiled plans with one or two levels per workgroup).
#pragma once
#include "fft_lds.h"

typedef MhGridArgs GridArgs;

namespace mh_grid {

constexpr int kMaxThreads = 768;
constexpr int kLdsBytes = 160 * 1024 - 256;

// FFT plan of the fused kernel for a half length Mh: radices in inverse order,
// the last (RL) being the register pass; TP = Mh/RL threads per level; RB
// levels per workgroup.  LDS: two buffers of RB rows of Mh+1 complex values,
// then the twiddles exp(-2 pi i k/Mh) and exp(-2 pi i k/(2Mh)), k < Mh.  Rows
// are Mh+2 values long (even, so every row starts on 16 bytes in FP32).
struct GPlan {
  int RL, TP, RB, nst, threads;
  int radix[16];
};

constexpr GPlan make_plan(int Mh) {
  GPlan p{};
  if (Mh < 2) return p;
  const int RL = Mh % 3 == 0 ? 3 : Mh % 4 == 0 ? 4 : Mh % 2 == 0 ? 2 : Mh % 5 == 0 ? 5 : 0;
  if (RL == 0) return p;
  int n = Mh / RL, k = 0;
  while (n % 4 == 0 && k < 15) { p.radix[k++] = 4; n /= 4; }
  while (n % 2 == 0 && k < 15) { p.radix[k++] = 2; n /= 2; }
  while (n % 3 == 0 && k < 15) { p.radix[k++] = 3; n /= 3; }
  while (n % 5 == 0 && k < 15) { p.radix[k++] = 5; n /= 5; }
  if (n != 1) return p;
  p.radix[k++] = RL;
  p.nst = k;
  p.TP = Mh / RL;
  for (int rb = 16; rb >= 1; --rb) {
    const int th = (rb * p.TP + 63) / 64 * 64;
    if (th <= kMaxThreads && (2 * rb * (Mh + 2) + 2 * Mh) * 16 <= kLdsBytes) {
      p.RB = rb;
      p.threads = th;
      break;
    }
  }
  if (p.RB > 0) p.RL = RL;
  return p;
}

// Half lengths with a compiled plan (all divisions and stage loops resolved
// at compile time); other lengths run the same code with the run-time plan.
#define MH_GRID_COMPILED(X) X(24) X(48) X(96) X(192) X(288) X(384) X(576) X(768)

// Workgroup barrier that waits for this wave's LDS operations only, so the
// global loads of the next component stay in flight across FFT passes.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

extern __shared__ double2 g_lds_raw[];

// Geometry of one launch; compile-time constants when MH > 0.
struct Geo {
  int Mh, M, RB, LS, TP, nst, threads;
  int RBL;  // rows per LDS buffer: RB levels times the components transformed together
};

template <int MH, int RBC, int NB = 1>
__device__ __forceinline__ Geo make_geo(const GridArgs& a, const GPlan& rp) {
  Geo g;
  g.M = a.M;
  if constexpr (MH > 0) {
    constexpr GPlan P = make_plan(MH);
    g.Mh = MH;
    g.RB = RBC > 0 ? RBC : P.RB;
    g.TP = P.TP;
    g.nst = P.nst;
    g.threads = RBC > 0 ? (RBC * P.TP + 63) / 64 * 64 : P.threads;
  } else {
    g.Mh = a.Mh;
    g.RB = rp.RB;
    g.TP = rp.TP;
    g.nst = rp.nst;
    g.threads = rp.threads;
  }
  g.LS = g.Mh + 2;
  g.RBL = g.RB * NB;
  return g;
}

// LDS: two buffers of RB rows of Mh+1 complex values, then the twiddles, all
// in the element type of the launch
template <typename T>
__device__ __forceinline__ typename vec2<T>::type* lds_base() {
  return reinterpret_cast<typename vec2<T>::type*>(g_lds_raw);
}
template <typename T>
__device__ __forceinline__ typename vec2<T>::type* lbuf(const Geo& g, int w) {
  return lds_base<T>() + w * g.RBL * g.LS;
}
template <typename T>
__device__ __forceinline__ const typename vec2<T>::type* ltw(const Geo& g) {
  return lds_base<T>() + 2 * g.RBL * g.LS;
}
template <typename T>
__device__ __forceinline__ const typename vec2<T>::type* ltw2(const Geo& g) {
  return ltw<T>(g) + g.Mh;
}

struct Ctx {
  const GPlan* rp;  // run-time plan (MH == 0)
  int row, q;       // register-pass mapping of this thread (row < RB when active)
  bool act;
  int busy;         // LDS buffer that other threads may still be reading
};

template <int R, typename T>
__device__ __forceinline__ void stage_mid(const Geo& g, const GridArgs& a, int win, int wout, int Ns,
                                          double dir, int nrows) {
  typedef typename vec2<T>::type V;
  const int nb = g.Mh / R, tstep = g.Mh / (Ns * R);
  const V* tw = ltw<T>(g);
  const V* in = lbuf<T>(g, win);
  V* out = lbuf<T>(g, wout);
  for (int idx = threadIdx.x; idx < nrows * nb; idx += g.threads) {
    const int row = idx / nb, bf = idx - row * nb;
    const int k = bf % Ns;
    const V* src = in + row * g.LS;
    V v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[bf + r * nb];
    if (Ns > 1) {
      // w^r by recurrence from one table value (one LDS read per butterfly)
      V w1 = tw[k * tstep];
      w1.y *= (T)-dir;  // tw holds exp(-i...); the inverse (dir = +1) conjugates
      V wr = w1;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        v[r] = cmul(v[r], wr);
        if (r + 1 < R) wr = cmul(wr, w1);
      }
    }
    dft<R>(v, dir);
    V* dst = out + row * g.LS + (bf / Ns) * Ns * R + k;
    if constexpr (sizeof(T) == 4 && R % 2 == 0) {
      if (Ns == 1) {
        // consecutive outputs from an even element: 16-byte stores of pairs
        // (half the store instructions at this pass's stride of R values)
#pragma unroll
        for (int r = 0; r < R; r += 2)
          *reinterpret_cast<float4*>(dst + r) = make_float4(v[r].x, v[r].y, v[r + 1].x, v[r + 1].y);
        continue;
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) dst[r * Ns] = v[r];
  }
}

template <typename T>
__device__ __forceinline__ void stage_mid_rt(const Geo& g, const GridArgs& a, int R, int win, int wout,
                                             int Ns, double dir, int nrows) {
  if (R == 4) stage_mid<4, T>(g, a, win, wout, Ns, dir, nrows);
  else if (R == 2) stage_mid<2, T>(g, a, win, wout, Ns, dir, nrows);
  else if (R == 3) stage_mid<3, T>(g, a, win, wout, Ns, dir, nrows);
  else stage_mid<5, T>(g, a, win, wout, Ns, dir, nrows);
}

// Inverse stages 0 .. nst-2 of a compiled plan (S = stage, NS = its Ns).
template <int MH, int S, int NS, typename T>
__device__ __forceinline__ int inv_mid_ct(const Geo& g, const GridArgs& a, int w, int nrows) {
  constexpr GPlan P = make_plan(MH);
  if constexpr (S < P.nst - 1) {
    constexpr int R = P.radix[S];
    stage_mid<R, T>(g, a, w, 1 - w, NS, 1.0, nrows);
    lds_barrier();
    return inv_mid_ct<MH, S + 1, NS * R, T>(g, a, 1 - w, nrows);
  } else {
    return w;
  }
}

// Forward stages S = nst-2 .. 0 of a compiled plan.
template <int MH, int S, int NS, typename T>
__device__ __forceinline__ int fwd_mid_ct(const Geo& g, const GridArgs& a, int w, int nrows) {
  constexpr GPlan P = make_plan(MH);
  if constexpr (S >= 0) {
    constexpr int R = P.radix[S];
    stage_mid<R, T>(g, a, w, 1 - w, NS, -1.0, nrows);
    lds_barrier();
    return fwd_mid_ct<MH, S - 1, NS * R, T>(g, a, 1 - w, nrows);
  } else {
    return w;
  }
}

// Global loads of one component's spectrum for the workgroup's levels, by
// conjugate pairs: item e = p*RB + j (pair p = 0 .. Mh/2, level j) holds X_p
// and X_{Mh-p} (zero where the order is >= M or the level is past the end),
// read from base + ((m*T + t)*Bst + b0 + j)*2 elements.  The offset is formed
// in 32 bits as m*(T*Bst*2) + ((t*Bst + b0 + j)*2) with a 24-bit multiply
// (full rate, where a 64-bit product costs four quarter-rate instructions);
// the launcher checks T*Bst*2 < 2^24 and that a component holds fewer than
// 2^31 elements, which makes both exact.
template <int RL, typename T>
struct Prefetch {
  static constexpr int PP = (RL + 2) / 2;  // items per thread (RL = 2, 3, 4)
  typename vec2<T>::type x[PP], xc[PP];
};

template <int MH, int RL, typename T>
__device__ __forceinline__ void fetch(const Geo& g, const GridArgs& a, Prefetch<RL, T>& pf,
                                      const T* base, int Bst, int b0, int nvalid, int t) {
  typedef typename vec2<T>::type V;
  const int RB = g.RB, n = (g.Mh / 2 + 1) * RB;
  const unsigned S = (unsigned)(a.T * Bst * 2);
  const unsigned tb = (unsigned)((t * Bst + b0) * 2);
  auto ld = [&](int m, int j) {
    return (m < g.M && j < nvalid)
               ? *reinterpret_cast<const V*>(base + ((unsigned)__umul24((unsigned)m, S) + tb +
                                                     2u * (unsigned)j))
               : mk2((T)0, (T)0);
  };
#pragma unroll
  for (int i = 0; i < Prefetch<RL, T>::PP; ++i) {
    const int e = threadIdx.x + i * g.threads;
    pf.x[i] = pf.xc[i] = mk2((T)0, (T)0);
    if (e < n) {
      const int p = e / RB, j = e - p * RB;
      pf.x[i] = ld(p, j);
      if (p > 0 && 2 * p < g.Mh) pf.xc[i] = ld(g.Mh - p, j);
    }
  }
}

struct NoFetch {
  __device__ void operator()() const {}
};

// Inverse transform of the prefetched component; returns this thread's RL
// complex grid values (x_{2p} + i x_{2p+1} at p = q + i*TP).  `next` issues
// the loads of the following component once the prefetched values are in
// LDS, so that they are in flight during this transform's passes.
template <int MH, int RL, typename T, typename F = NoFetch>
__device__ __forceinline__ void inverse(Ctx& c, const Geo& g, const GridArgs& a,
                                        const Prefetch<RL, T>& pf, typename vec2<T>::type (&z)[RL],
                                        F next = F()) {
  typedef typename vec2<T>::type V;
  const int RB = g.RB, LS = g.LS;
  int w = 1 - c.busy;
  {
    // packed spectrum of the real inverse FFT, written by conjugate pairs:
    // Z_k = E + iO, E = X_k + conj(X_{Mh-k}), O = (X_k - conj(X_{Mh-k})) e^{+2 pi i k/n},
    // and Z_{Mh-k} = conj(E) + i conj(O) (X_0 is real: its imaginary part is dropped)
    V* zb = lbuf<T>(g, w);
    const V* tw2 = ltw2<T>(g);
    const int n = (g.Mh / 2 + 1) * RB;
#pragma unroll
    for (int i = 0; i < Prefetch<RL, T>::PP; ++i) {
      const int e = threadIdx.x + i * g.threads;
      if (e < n) {
        const int p = e / RB, j = e - p * RB;
        V x = pf.x[i];
        if (p == 0) x.y = (T)0;
        const V xc = (2 * p == g.Mh) ? cconj(x) : cconj(pf.xc[i]);
        const V E = cadd(x, xc);
        const V O = cmul(csub(x, xc), cconj(tw2[p]));
        V* row = zb + j * LS;
        row[p] = mk2(E.x - O.y, E.y + O.x);
        if (p > 0 && 2 * p < g.Mh) row[g.Mh - p] = mk2(E.x + O.y, -E.y + O.x);
      }
    }
  }
  next();
  lds_barrier();
  if (g.nst > 1) {
    if constexpr (MH > 0) {
      w = inv_mid_ct<MH, 0, 1, T>(g, a, w, g.RB);
    } else {
      const GPlan& P = *c.rp;
      int Ns = 1;
      for (int s = 0; s < P.nst - 1; ++s) {
        stage_mid_rt<T>(g, a, P.radix[s], w, 1 - w, Ns, 1.0, g.RB);
        Ns *= P.radix[s];
        w = 1 - w;
        lds_barrier();
      }
    }
  }
  // last pass (radix RL, Ns = TP): butterfly q of level `row`, twiddles w^{i q}
  if (c.act) {
    const V* src = lbuf<T>(g, w) + c.row * LS;
    const V* tw = ltw<T>(g);
#pragma unroll
    for (int i = 0; i < RL; ++i) z[i] = src[c.q + i * g.TP];
    if (g.nst > 1) {
      const V w1 = cconj(tw[c.q]);
      V wi = w1;
#pragma unroll
      for (int i = 1; i < RL; ++i) {
        z[i] = cmul(z[i], wi);
        if (i + 1 < RL) wi = cmul(wi, w1);
      }
    }
    dft<RL>(z, 1.0);
  } else {
#pragma unroll
    for (int i = 0; i < RL; ++i) z[i] = mk2((T)0, (T)0);
  }
  c.busy = w;
}

// Forward transform of the thread's grid values z; writes X_m (m < M) times
// `scale` into output component `comp` of Fo for the bulk levels of the block.
template <int MH, int RL, typename T>
__device__ __forceinline__ void forward(Ctx& c, const Geo& g, const GridArgs& a,
                                        typename vec2<T>::type (&z)[RL], int comp, T scale, int t,
                                        int r0) {
  typedef typename vec2<T>::type V;
  const int LS = g.LS;
  int w = 1 - c.busy;
  // first pass (radix RL, Ns = 1): inputs at q + i*TP, outputs at q*RL + i
  if (c.act) {
    dft<RL>(z, -1.0);
    V* dst = lbuf<T>(g, w) + c.row * LS + c.q * RL;
#pragma unroll
    for (int i = 0; i < RL; ++i) dst[i] = z[i];
  }
  lds_barrier();
  if constexpr (MH > 0) {
    constexpr GPlan P = make_plan(MH);
    w = fwd_mid_ct<MH, P.nst - 2, RL, T>(g, a, w, g.RB);
  } else {
    const GPlan& P = *c.rp;
    int Ns = RL;
    for (int s = P.nst - 2; s >= 0; --s) {
      stage_mid_rt<T>(g, a, P.radix[s], w, 1 - w, Ns, -1.0, g.RB);
      Ns *= P.radix[s];
      w = 1 - w;
      lds_barrier();
    }
  }
  // unpack by conjugate pairs: X_m = E + e^{-2 pi i m/n} O and
  // X_{Mh-m} = conj(E - e^{-2 pi i m/n} O), E = (Z_m + conj Z_{Mh-m})/2,
  // O = (Z_m - conj Z_{Mh-m})/(2i); store the orders < M of the bulk levels
  const int RB = g.RB, Mh = g.Mh, N = a.N, No = N - 2, n = (Mh / 2 + 1) * RB;
  const V* tw2 = ltw2<T>(g);
  T* Fo = reinterpret_cast<T*>(a.Fo) + (long)comp * g.M * a.T * No * 2;
  // 32-bit offsets as in fetch(): p*(T*No*2) + (t*No + r - 1)*2
  const unsigned So = (unsigned)(a.T * No * 2);
  const unsigned to = (unsigned)((t * No + r0 - 1) * 2);
  const V* res = lbuf<T>(g, w);
  for (int e = threadIdx.x; e < n; e += g.threads) {
    const int p = e / RB, j = e - p * RB;
    const int r = r0 + j;
    if (r < 1 || r > N - 2) continue;
    const V* src = res + j * LS;
    const int pc = (p == 0) ? 0 : Mh - p;
    const V Z = src[p];
    const V Zc = cconj(src[pc]);
    const V E = cscale(cadd(Z, Zc), (T)0.5);
    const V D = csub(Z, Zc);
    const V O = mk2((T)0.5 * D.y, (T)-0.5 * D.x);
    const V Tm = cmul(tw2[p], O);
    const unsigned oj = to + 2u * (unsigned)j;
    if (p < g.M)
      *reinterpret_cast<V*>(Fo + ((unsigned)__umul24((unsigned)p, So) + oj)) =
          cscale(cadd(E, Tm), scale);
    if (p > 0 && 2 * p < Mh && pc < g.M)
      *reinterpret_cast<V*>(Fo + ((unsigned)__umul24((unsigned)pc, So) + oj)) =
          cscale(cconj(csub(E, Tm)), scale);
  }
  c.busy = w;
}

// ---------------------------------------------------------------- batched transforms
// The FP32 kernel of the compiled plans transforms the three components of a
// vector (u, B, curl u, curl B; the three outputs of each product) together:
// component k of level j is LDS row k*RB + j, and every LDS pass runs over the
// 3*RB rows, so a vector costs the barriers of one transform instead of three.
// Each row's arithmetic is that of the single-component transform.

template <int RL, typename T, int NC>
struct PrefetchN {
  static constexpr int PP = (RL + 2) / 2;
  typename vec2<T>::type x[NC][PP], xc[NC][PP];
};

// Loads of NC components at base + k*cs (k < NC), as fetch().
template <int MH, int RL, typename T, int NC>
__device__ __forceinline__ void fetchN(const Geo& g, const GridArgs& a, PrefetchN<RL, T, NC>& pf,
                                       const T* base, long cs, int Bst, int b0, int nvalid, int t) {
  typedef typename vec2<T>::type V;
  const int RB = g.RB, n = (g.Mh / 2 + 1) * RB;
  const unsigned S = (unsigned)(a.T * Bst * 2);
  const unsigned tb = (unsigned)((t * Bst + b0) * 2);
  auto ld = [&](const T* b, int m, int j) {
    return (m < g.M && j < nvalid)
               ? *reinterpret_cast<const V*>(b + ((unsigned)__umul24((unsigned)m, S) + tb +
                                                  2u * (unsigned)j))
               : mk2((T)0, (T)0);
  };
#pragma unroll
  for (int i = 0; i < PrefetchN<RL, T, NC>::PP; ++i) {
    const int e = threadIdx.x + i * g.threads;
#pragma unroll
    for (int k = 0; k < NC; ++k) pf.x[k][i] = pf.xc[k][i] = mk2((T)0, (T)0);
    if (e < n) {
      const int p = e / RB, j = e - p * RB;
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        pf.x[k][i] = ld(base + k * cs, p, j);
        if (p > 0 && 2 * p < g.Mh) pf.xc[k][i] = ld(base + k * cs, g.Mh - p, j);
      }
    }
  }
}

// Inverse transforms of NC prefetched components (as inverse()).
template <int MH, int RL, typename T, int NC, typename F = NoFetch>
__device__ __forceinline__ void inverseN(Ctx& c, const Geo& g, const GridArgs& a,
                                         const PrefetchN<RL, T, NC>& pf,
                                         typename vec2<T>::type (&z)[NC][RL], F next = F()) {
  typedef typename vec2<T>::type V;
  const int RB = g.RB, LS = g.LS;
  int w = 1 - c.busy;
  {
    V* zb = lbuf<T>(g, w);
    const V* tw2 = ltw2<T>(g);
    const int n = (g.Mh / 2 + 1) * RB;
#pragma unroll
    for (int i = 0; i < PrefetchN<RL, T, NC>::PP; ++i) {
      const int e = threadIdx.x + i * g.threads;
      if (e < n) {
        const int p = e / RB, j = e - p * RB;
        const V tw2c = cconj(tw2[p]);
#pragma unroll
        for (int k = 0; k < NC; ++k) {
          V x = pf.x[k][i];
          if (p == 0) x.y = (T)0;
          const V xc = (2 * p == g.Mh) ? cconj(x) : cconj(pf.xc[k][i]);
          const V E = cadd(x, xc);
          const V O = cmul(csub(x, xc), tw2c);
          V* row = zb + (k * RB + j) * LS;
          row[p] = mk2(E.x - O.y, E.y + O.x);
          if (p > 0 && 2 * p < g.Mh) row[g.Mh - p] = mk2(E.x + O.y, -E.y + O.x);
        }
      }
    }
  }
  next();
  lds_barrier();
  if (g.nst > 1) w = inv_mid_ct<MH, 0, 1, T>(g, a, w, NC * RB);
  if (c.act) {
    const V* tw = ltw<T>(g);
    const V w1 = cconj(tw[c.q]);
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const V* src = lbuf<T>(g, w) + (k * RB + c.row) * LS;
#pragma unroll
      for (int i = 0; i < RL; ++i) z[k][i] = src[c.q + i * g.TP];
      if (g.nst > 1) {
        V wi = w1;
#pragma unroll
        for (int i = 1; i < RL; ++i) {
          z[k][i] = cmul(z[k][i], wi);
          if (i + 1 < RL) wi = cmul(wi, w1);
        }
      }
      dft<RL>(z[k], 1.0);
    }
  } else {
#pragma unroll
    for (int k = 0; k < NC; ++k)
#pragma unroll
      for (int i = 0; i < RL; ++i) z[k][i] = mk2((T)0, (T)0);
  }
  c.busy = w;
}

// Forward transforms of NC components (as forward()); output component
// comp[k] scaled by scale[k].
template <int MH, int RL, typename T, int NC>
__device__ __forceinline__ void forwardN(Ctx& c, const Geo& g, const GridArgs& a,
                                         typename vec2<T>::type (&z)[NC][RL], const int (&comp)[NC],
                                         const T (&scale)[NC], int t, int r0) {
  typedef typename vec2<T>::type V;
  const int LS = g.LS, RB = g.RB;
  int w = 1 - c.busy;
  if (c.act) {
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      dft<RL>(z[k], -1.0);
      V* dst = lbuf<T>(g, w) + (k * RB + c.row) * LS + c.q * RL;
#pragma unroll
      for (int i = 0; i < RL; ++i) dst[i] = z[k][i];
    }
  }
  lds_barrier();
  {
    constexpr GPlan P = make_plan(MH);
    w = fwd_mid_ct<MH, P.nst - 2, RL, T>(g, a, w, NC * RB);
  }
  const int Mh = g.Mh, N = a.N, No = N - 2, n = (Mh / 2 + 1) * RB;
  const V* tw2 = ltw2<T>(g);
  const unsigned So = (unsigned)(a.T * No * 2);
  const unsigned to = (unsigned)((t * No + r0 - 1) * 2);
  const V* res = lbuf<T>(g, w);
  for (int e = threadIdx.x; e < n; e += g.threads) {
    const int p = e / RB, j = e - p * RB;
    const int r = r0 + j;
    if (r < 1 || r > N - 2) continue;
    const int pc = (p == 0) ? 0 : Mh - p;
    const unsigned oj = to + 2u * (unsigned)j;
    const V tw = tw2[p];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const V* src = res + (k * RB + j) * LS;
      T* Fo = reinterpret_cast<T*>(a.Fo) + (long)comp[k] * g.M * a.T * No * 2;
      const V Z = src[p];
      const V Zc = cconj(src[pc]);
      const V E = cscale(cadd(Z, Zc), (T)0.5);
      const V D = csub(Z, Zc);
      const V O = mk2((T)0.5 * D.y, (T)-0.5 * D.x);
      const V Tm = cmul(tw, O);
      if (p < g.M)
        *reinterpret_cast<V*>(Fo + ((unsigned)__umul24((unsigned)p, So) + oj)) =
            cscale(cadd(E, Tm), scale[k]);
      if (p > 0 && 2 * p < Mh && pc < g.M)
        *reinterpret_cast<V*>(Fo + ((unsigned)__umul24((unsigned)pc, So) + oj)) =
            cscale(cconj(csub(E, Tm)), scale[k]);
    }
  }
  c.busy = w;
}

}  // namespace mh_grid
using namespace mh_grid;

template <int MH, int RL, typename T, int RBC>
__global__ __launch_bounds__(kMaxThreads, 3) void k_grid_step(GridArgs a, GPlan rp) {
  typedef typename vec2<T>::type V;
  __shared__ unsigned long long red[2 * 16];
  const Geo g = make_geo<MH, RBC>(a, rp);
  const int N = a.N, nth = a.T, RB = g.RB;
  const int nRB = (N + RB - 1) / RB;
  // XCD-aware order: workgroups i and i+8 are dispatched to the same XCD, so
  // consecutive logical indices (neighbouring level blocks of one theta row,
  // which share 128-byte lines of F) are given to them.
  const int nwg = nth * nRB, per = (nwg + 7) / 8;
  const int L = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (L >= nwg) return;  // uniform per workgroup
  const int t = L / nRB, r0 = (L - t * nRB) * RB;
  const int nvalid = min(RB, N - r0);

  Ctx c;
  c.rp = &rp;
  c.row = threadIdx.x / g.TP;
  c.q = threadIdx.x - c.row * g.TP;
  c.act = c.row < nvalid;
  c.busy = 0;
  if (threadIdx.x < 2 * RB) red[threadIdx.x] = 0ull;
  {  // twiddles into LDS
    V* tl = lds_base<T>() + 2 * RB * g.LS;
    const double2* t1 = reinterpret_cast<const double2*>(a.tw);  // FP64 tables
    const double2* t2 = reinterpret_cast<const double2*>(a.tw2);
    for (int k = threadIdx.x; k < g.Mh; k += g.threads) {
      tl[k] = mk2((T)t1[k].x, (T)t1[k].y);
      tl[g.Mh + k] = mk2((T)t2[k].x, (T)t2[k].y);
    }
  }
  // the first inverse reads tw2 from LDS while it writes its packed spectrum,
  // before its own first barrier: every twiddle (and the zeroed maxima) must
  // be in place first
  __syncthreads();

  const int r = r0 + (c.act ? c.row : 0);
  const T ost2 = (T)a.ost2[t];
  const T or2 = (T)a.or2[r], or4 = (T)a.or4[r];
  const T LF = (T)a.LFfac;
  const int Bv = 4 * N;
  const long vcs = (long)g.M * a.T * Bv * 2;  // component stride of Fv (elements)
  const T* Fv = reinterpret_cast<const T*>(a.Fv);
  const T* Fs = reinterpret_cast<const T*>(a.Fs);
  const bool mag = a.l_mag != 0;

  Prefetch<RL, T> pf;
  // loads of a component, issued by inverse() after it has packed the last one
  auto next = [&](const T* base, int Bst, int b0) {
    return [&, base, Bst, b0]() { fetch<MH, RL, T>(g, a, pf, base, Bst, b0, nvalid, t); };
  };
  V vr[RL], vt[RL], vp[RL], tmp[RL];
  const T inv_n = (T)(1.0 / (2.0 * g.Mh));
  // 1. u (group 2)
  fetch<MH, RL, T>(g, a, pf, Fv, Bv, 2 * N + r0, nvalid, t);
  inverse<MH, RL, T>(c, g, a, pf, vr, next(Fv + vcs, Bv, 2 * N + r0));
  inverse<MH, RL, T>(c, g, a, pf, vt, next(Fv + 2 * vcs, Bv, 2 * N + r0));
  inverse<MH, RL, T>(c, g, a, pf, vp, next(Fs, N, r0));
  // 2. s -> VSr, VSt, VSp
  // B_r next, or curl u_r without a magnetic field
  inverse<MH, RL, T>(c, g, a, pf, tmp, next(Fv, Bv, (mag ? 0 : 3 * N) + r0));
  {
    V o[RL];
#pragma unroll
    for (int i = 0; i < RL; ++i) o[i] = mk2(vr[i].x * tmp[i].x, vr[i].y * tmp[i].y);
    forward<MH, RL, T>(c, g, a, o, 1, inv_n, t, r0);
#pragma unroll
    for (int i = 0; i < RL; ++i)
      o[i] = mk2(or2 * vt[i].x * tmp[i].x, or2 * vt[i].y * tmp[i].y);
    forward<MH, RL, T>(c, g, a, o, 5, inv_n * ost2, t, r0);
#pragma unroll
    for (int i = 0; i < RL; ++i)
      o[i] = mk2(or2 * vp[i].x * tmp[i].x, or2 * vp[i].y * tmp[i].y);
    forward<MH, RL, T>(c, g, a, o, 6, inv_n * ost2, t, r0);
  }
  // 3. B -> Courant maxima (all levels incl. boundaries) and VxBr, VxBt, VxBp.
  // In FP32 the B values stay in registers for the Lorentz force below; in
  // FP64 they are transformed again there (register budget).
  constexpr bool keepB = sizeof(T) == 4;
  V br[RL], bt[RL], bp[RL];
  {
    if (mag) {
      inverse<MH, RL, T>(c, g, a, pf, br, next(Fv + vcs, Bv, r0));
      inverse<MH, RL, T>(c, g, a, pf, bt, next(Fv + 2 * vcs, Bv, r0));
      inverse<MH, RL, T>(c, g, a, pf, bp, next(Fv, Bv, 3 * N + r0));  // curl u_r next
    } else {
#pragma unroll
      for (int i = 0; i < RL; ++i) br[i] = bt[i] = bp[i] = mk2((T)0, (T)0);
    }
    if (c.act) {
      T cr = (T)0, ch = (T)0;
      const T vri = (T)a.valri2[r], vhi = (T)a.valhi2[r];
#pragma unroll
      for (int i = 0; i < RL; ++i) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const T Vr = h ? vr[i].y : vr[i].x, Vt = h ? vt[i].y : vt[i].x;
          const T Vp = h ? vp[i].y : vp[i].x, Br = h ? br[i].y : br[i].x;
          const T Bt = h ? bt[i].y : bt[i].x, Bp = h ? bp[i].y : bp[i].x;
          const T vflr2 = Vr * Vr;
          const T vflh2 = (Vt * Vt + Vp * Vp) * ost2;
          T c1, c2;
          if (mag) {
            const T valr = Br * Br * LF;
            const T valh2 = (Bt * Bt + Bp * Bp) * LF * ost2;
            T valr2 = valr, valh2m = valh2;
            if (a.alf_damp) {
              valr2 = valr * valr / (valr + vri);
              valh2m = valh2 * valh2 / (valh2 + vhi);
            }
            c1 = or4 * ((T)a.cf2 * vflr2 + (T)a.af2 * valr2);
            c2 = or2 * ((T)a.cf2 * vflh2 + (T)a.af2 * valh2m);
          } else {
            c1 = (T)a.cf2 * or4 * vflr2;
            c2 = (T)a.cf2 * or2 * vflh2;
          }
          cr = fmax(cr, c1);
          ch = fmax(ch, c2);
        }
      }
      atomicMax(&red[c.row], (unsigned long long)__double_as_longlong((double)cr));
      atomicMax(&red[RB + c.row], (unsigned long long)__double_as_longlong((double)ch));
    }
    V o[RL];
#pragma unroll
    for (int i = 0; i < RL; ++i)
      o[i] = mk2(ost2 * (vt[i].x * bp[i].x - vp[i].x * bt[i].x),
                          ost2 * (vt[i].y * bp[i].y - vp[i].y * bt[i].y));
    forward<MH, RL, T>(c, g, a, o, 2, inv_n, t, r0);
#pragma unroll
    for (int i = 0; i < RL; ++i)
      o[i] = mk2(or4 * (vp[i].x * br[i].x - vr[i].x * bp[i].x),
                          or4 * (vp[i].y * br[i].y - vr[i].y * bp[i].y));
    forward<MH, RL, T>(c, g, a, o, 7, inv_n * ost2, t, r0);
#pragma unroll
    for (int i = 0; i < RL; ++i)
      o[i] = mk2(or4 * (vr[i].x * bt[i].x - vt[i].x * br[i].x),
                          or4 * (vr[i].y * bt[i].y - vt[i].y * br[i].y));
    forward<MH, RL, T>(c, g, a, o, 8, inv_n * ost2, t, r0);
  }
  // 4-5. Advection + Lorentz force, accumulated one component of curl u / curl B
  // at a time (B is transformed a second time so that u and B are never live
  // together with the accumulators):
  //   Advr = -ost2 (Cvt Vp - Cvp Vt) + LF ost2 (Cbt Bp - Cbp Bt)
  //   Advt = -or4 (Cvp Vr - Cvr Vp)  + LF or4 (Cbp Br - Cbr Bp)
  //   Advp = -or4 (Cvr Vt - Cvt Vr)  + LF or4 (Cbr Bt - Cbt Br)
  V ar[RL], at[RL], ap[RL];
  inverse<MH, RL, T>(c, g, a, pf, tmp, next(Fv + vcs, Bv, 3 * N + r0));  // curl u_r
#pragma unroll
  for (int i = 0; i < RL; ++i) {
    ar[i] = mk2((T)0, (T)0);
    at[i] = mk2(or4 * tmp[i].x * vp[i].x, or4 * tmp[i].y * vp[i].y);
    ap[i] = mk2(-or4 * tmp[i].x * vt[i].x, -or4 * tmp[i].y * vt[i].y);
  }
  inverse<MH, RL, T>(c, g, a, pf, tmp, next(Fv + 2 * vcs, Bv, 3 * N + r0));  // curl u_t
#pragma unroll
  for (int i = 0; i < RL; ++i) {
    ar[i].x -= ost2 * tmp[i].x * vp[i].x;
    ar[i].y -= ost2 * tmp[i].y * vp[i].y;
    ap[i].x += or4 * tmp[i].x * vr[i].x;
    ap[i].y += or4 * tmp[i].y * vr[i].y;
  }
  // curl u_p; then curl B_r (B kept in registers) or B_r (B transformed again)
  inverse<MH, RL, T>(c, g, a, pf, tmp, [&]() {
    if (mag) fetch<MH, RL, T>(g, a, pf, Fv, Bv, (keepB ? N : 0) + r0, nvalid, t);
  });
#pragma unroll
  for (int i = 0; i < RL; ++i) {
    ar[i].x += ost2 * tmp[i].x * vt[i].x;
    ar[i].y += ost2 * tmp[i].y * vt[i].y;
    at[i].x -= or4 * tmp[i].x * vr[i].x;
    at[i].y -= or4 * tmp[i].y * vr[i].y;
  }
  if (mag) {
    const T lr = LF * ost2, lt = LF * or4;
    if (!keepB) {
      inverse<MH, RL, T>(c, g, a, pf, br, next(Fv + vcs, Bv, r0));
      inverse<MH, RL, T>(c, g, a, pf, bt, next(Fv + 2 * vcs, Bv, r0));
      inverse<MH, RL, T>(c, g, a, pf, bp, next(Fv, Bv, N + r0));
    }
    inverse<MH, RL, T>(c, g, a, pf, tmp, next(Fv + vcs, Bv, N + r0));  // curl B_r
#pragma unroll
    for (int i = 0; i < RL; ++i) {
      at[i].x -= lt * tmp[i].x * bp[i].x;
      at[i].y -= lt * tmp[i].y * bp[i].y;
      ap[i].x += lt * tmp[i].x * bt[i].x;
      ap[i].y += lt * tmp[i].y * bt[i].y;
    }
    inverse<MH, RL, T>(c, g, a, pf, tmp, next(Fv + 2 * vcs, Bv, N + r0));  // curl B_t
#pragma unroll
    for (int i = 0; i < RL; ++i) {
      ar[i].x += lr * tmp[i].x * bp[i].x;
      ar[i].y += lr * tmp[i].y * bp[i].y;
      ap[i].x -= lt * tmp[i].x * br[i].x;
      ap[i].y -= lt * tmp[i].y * br[i].y;
    }
    inverse<MH, RL, T>(c, g, a, pf, tmp);  // curl B_p
#pragma unroll
    for (int i = 0; i < RL; ++i) {
      ar[i].x -= lr * tmp[i].x * bt[i].x;
      ar[i].y -= lr * tmp[i].y * bt[i].y;
      at[i].x += lt * tmp[i].x * br[i].x;
      at[i].y += lt * tmp[i].y * br[i].y;
    }
  }
  forward<MH, RL, T>(c, g, a, ar, 0, inv_n, t, r0);
  forward<MH, RL, T>(c, g, a, at, 3, inv_n * ost2, t, r0);
  forward<MH, RL, T>(c, g, a, ap, 4, inv_n * ost2, t, r0);
  lds_barrier();
  if (threadIdx.x < 2 * nvalid) {
    const int j = threadIdx.x % nvalid, h = threadIdx.x / nvalid;
    atomicMax(a.cfl + h * N + r0 + j, red[h * RB + j]);
  }
}

// FP32 grid step of the compiled plans with the vector components transformed
// together (inverseN / forwardN over 3*RB LDS rows).  Same arithmetic per
// value as k_grid_step<MH, RL, float, RBC>, in the same order.
template <int MH, int RL, int RBC>
__global__ __launch_bounds__(kMaxThreads, 3) void k_grid_step3(GridArgs a, GPlan rp) {
  typedef float T;
  typedef typename vec2<T>::type V;
  __shared__ unsigned long long red[2 * 16];
  const Geo g = make_geo<MH, RBC, 3>(a, rp);
  const int N = a.N, nth = a.T, RB = g.RB;
  const int nRB = (N + RB - 1) / RB;
  const int nwg = nth * nRB, per = (nwg + 7) / 8;
  const int L = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
  if (L >= nwg) return;  // uniform per workgroup
  const int t = L / nRB, r0 = (L - t * nRB) * RB;
  const int nvalid = min(RB, N - r0);

  Ctx c;
  c.rp = &rp;
  c.row = threadIdx.x / g.TP;
  c.q = threadIdx.x - c.row * g.TP;
  c.act = c.row < nvalid;
  c.busy = 0;
  if (threadIdx.x < 2 * RB) red[threadIdx.x] = 0ull;
  {  // twiddles into LDS
    V* tl = lds_base<T>() + 2 * g.RBL * g.LS;
    const double2* t1 = reinterpret_cast<const double2*>(a.tw);
    const double2* t2 = reinterpret_cast<const double2*>(a.tw2);
    for (int k = threadIdx.x; k < g.Mh; k += g.threads) {
      tl[k] = mk2((T)t1[k].x, (T)t1[k].y);
      tl[g.Mh + k] = mk2((T)t2[k].x, (T)t2[k].y);
    }
  }
  __syncthreads();

  const int r = r0 + (c.act ? c.row : 0);
  const T ost2 = (T)a.ost2[t];
  const T or2 = (T)a.or2[r], or4 = (T)a.or4[r];
  const T LF = (T)a.LFfac;
  const int Bv = 4 * N;
  const long vcs = (long)g.M * a.T * Bv * 2;
  const T* Fv = reinterpret_cast<const T*>(a.Fv);
  const T* Fs = reinterpret_cast<const T*>(a.Fs);
  const bool mag = a.l_mag != 0;
  const T inv_n = (T)(1.0 / (2.0 * g.Mh));

  PrefetchN<RL, T, 3> pf3;
  PrefetchN<RL, T, 1> pf1;
  auto next3 = [&](int b0) {
    return [&, b0]() { fetchN<MH, RL, T, 3>(g, a, pf3, Fv, vcs, Bv, b0, nvalid, t); };
  };
  V u[3][RL];  // vr, vt, vp
  // 1. u (group 2), then s
  fetchN<MH, RL, T, 3>(g, a, pf3, Fv, vcs, Bv, 2 * N + r0, nvalid, t);
  inverseN<MH, RL, T, 3>(c, g, a, pf3, u,
                         [&]() { fetchN<MH, RL, T, 1>(g, a, pf1, Fs, 0, N, r0, nvalid, t); });
  // 2. s -> VSr, VSt, VSp; B next, or curl u without a magnetic field
  {
    V sv[1][RL];
    inverseN<MH, RL, T, 1>(c, g, a, pf1, sv, next3((mag ? 0 : 3 * N) + r0));
    V o[3][RL];
#pragma unroll
    for (int i = 0; i < RL; ++i) {
      o[0][i] = mk2(u[0][i].x * sv[0][i].x, u[0][i].y * sv[0][i].y);
      o[1][i] = mk2(or2 * u[1][i].x * sv[0][i].x, or2 * u[1][i].y * sv[0][i].y);
      o[2][i] = mk2(or2 * u[2][i].x * sv[0][i].x, or2 * u[2][i].y * sv[0][i].y);
    }
    const int comp[3] = {1, 5, 6};
    const T sc[3] = {inv_n, inv_n * ost2, inv_n * ost2};
    forwardN<MH, RL, T, 3>(c, g, a, o, comp, sc, t, r0);
  }
  // 3. B -> Courant maxima and VxBr, VxBt, VxBp (B stays in registers)
  V b[3][RL];
  if (mag) {
    inverseN<MH, RL, T, 3>(c, g, a, pf3, b, next3(3 * N + r0));  // curl u next
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int i = 0; i < RL; ++i) b[k][i] = mk2((T)0, (T)0);
  }
  if (c.act) {
    T cr = (T)0, ch = (T)0;
    const T vri = (T)a.valri2[r], vhi = (T)a.valhi2[r];
#pragma unroll
    for (int i = 0; i < RL; ++i) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const T Vr = h ? u[0][i].y : u[0][i].x, Vt = h ? u[1][i].y : u[1][i].x;
        const T Vp = h ? u[2][i].y : u[2][i].x, Br = h ? b[0][i].y : b[0][i].x;
        const T Bt = h ? b[1][i].y : b[1][i].x, Bp = h ? b[2][i].y : b[2][i].x;
        const T vflr2 = Vr * Vr;
        const T vflh2 = (Vt * Vt + Vp * Vp) * ost2;
        T c1, c2;
        if (mag) {
          const T valr = Br * Br * LF;
          const T valh2 = (Bt * Bt + Bp * Bp) * LF * ost2;
          T valr2 = valr, valh2m = valh2;
          if (a.alf_damp) {
            valr2 = valr * valr / (valr + vri);
            valh2m = valh2 * valh2 / (valh2 + vhi);
          }
          c1 = or4 * ((T)a.cf2 * vflr2 + (T)a.af2 * valr2);
          c2 = or2 * ((T)a.cf2 * vflh2 + (T)a.af2 * valh2m);
        } else {
          c1 = (T)a.cf2 * or4 * vflr2;
          c2 = (T)a.cf2 * or2 * vflh2;
        }
        cr = fmax(cr, c1);
        ch = fmax(ch, c2);
      }
    }
    atomicMax(&red[c.row], (unsigned long long)__double_as_longlong((double)cr));
    atomicMax(&red[RB + c.row], (unsigned long long)__double_as_longlong((double)ch));
  }
  {
    V o[3][RL];
#pragma unroll
    for (int i = 0; i < RL; ++i) {
      o[0][i] = mk2(ost2 * (u[1][i].x * b[2][i].x - u[2][i].x * b[1][i].x),
                    ost2 * (u[1][i].y * b[2][i].y - u[2][i].y * b[1][i].y));
      o[1][i] = mk2(or4 * (u[2][i].x * b[0][i].x - u[0][i].x * b[2][i].x),
                    or4 * (u[2][i].y * b[0][i].y - u[0][i].y * b[2][i].y));
      o[2][i] = mk2(or4 * (u[0][i].x * b[1][i].x - u[1][i].x * b[0][i].x),
                    or4 * (u[0][i].y * b[1][i].y - u[1][i].y * b[0][i].y));
    }
    const int comp[3] = {2, 7, 8};
    const T sc[3] = {inv_n, inv_n * ost2, inv_n * ost2};
    forwardN<MH, RL, T, 3>(c, g, a, o, comp, sc, t, r0);
  }
  // 4-5. Advection + Lorentz force (formulas as k_grid_step)
  V ad[3][RL];  // Advr, Advt, Advp
  {
    V cu[3][RL];  // curl u
    inverseN<MH, RL, T, 3>(c, g, a, pf3, cu, [&]() {
      if (mag) fetchN<MH, RL, T, 3>(g, a, pf3, Fv, vcs, Bv, N + r0, nvalid, t);
    });
#pragma unroll
    for (int i = 0; i < RL; ++i) {
      ad[0][i] = mk2((T)0, (T)0);
      ad[1][i] = mk2(or4 * cu[0][i].x * u[2][i].x, or4 * cu[0][i].y * u[2][i].y);
      ad[2][i] = mk2(-or4 * cu[0][i].x * u[1][i].x, -or4 * cu[0][i].y * u[1][i].y);
    }
#pragma unroll
    for (int i = 0; i < RL; ++i) {
      ad[0][i].x -= ost2 * cu[1][i].x * u[2][i].x;
      ad[0][i].y -= ost2 * cu[1][i].y * u[2][i].y;
      ad[2][i].x += or4 * cu[1][i].x * u[0][i].x;
      ad[2][i].y += or4 * cu[1][i].y * u[0][i].y;
    }
#pragma unroll
    for (int i = 0; i < RL; ++i) {
      ad[0][i].x += ost2 * cu[2][i].x * u[1][i].x;
      ad[0][i].y += ost2 * cu[2][i].y * u[1][i].y;
      ad[1][i].x -= or4 * cu[2][i].x * u[0][i].x;
      ad[1][i].y -= or4 * cu[2][i].y * u[0][i].y;
    }
  }
  if (mag) {
    const T lr = LF * ost2, lt = LF * or4;
    V cb[3][RL];  // curl B
    inverseN<MH, RL, T, 3>(c, g, a, pf3, cb);
#pragma unroll
    for (int i = 0; i < RL; ++i) {
      ad[1][i].x -= lt * cb[0][i].x * b[2][i].x;
      ad[1][i].y -= lt * cb[0][i].y * b[2][i].y;
      ad[2][i].x += lt * cb[0][i].x * b[1][i].x;
      ad[2][i].y += lt * cb[0][i].y * b[1][i].y;
    }
#pragma unroll
    for (int i = 0; i < RL; ++i) {
      ad[0][i].x += lr * cb[1][i].x * b[2][i].x;
      ad[0][i].y += lr * cb[1][i].y * b[2][i].y;
      ad[2][i].x -= lt * cb[1][i].x * b[0][i].x;
      ad[2][i].y -= lt * cb[1][i].y * b[0][i].y;
    }
#pragma unroll
    for (int i = 0; i < RL; ++i) {
      ad[0][i].x -= lr * cb[2][i].x * b[1][i].x;
      ad[0][i].y -= lr * cb[2][i].y * b[1][i].y;
      ad[1][i].x += lt * cb[2][i].x * b[0][i].x;
      ad[1][i].y += lt * cb[2][i].y * b[0][i].y;
    }
  }
  {
    const int comp[3] = {0, 3, 4};
    const T sc[3] = {inv_n, inv_n * ost2, inv_n * ost2};
    forwardN<MH, RL, T, 3>(c, g, a, ad, comp, sc, t, r0);
  }
  lds_barrier();
  if (threadIdx.x < 2 * nvalid) {
    const int j = threadIdx.x % nvalid, h = threadIdx.x / nvalid;
    atomicMax(a.cfl + h * N + r0 + j, red[h * RB + j]);
  }
}

// Compiled plan for the common half lengths, run-time plan otherwise.

// Launch of a compiled plan with RBC levels per workgroup (0: the plan's own);
// false when Mh has no compiled plan.
template <int RBC, typename T>
bool launch_compiled(const GridArgs& a, const GPlan& P, dim3 gd, dim3 bd, size_t lds,
                     hipStream_t st) {
  switch (a.Mh) {
#define MH_CASE(V)                                                                       \
  case V:                                                                                \
    if constexpr (sizeof(T) == 4)                                                        \
      hipLaunchKernelGGL((k_grid_step3<V, make_plan(V).RL, RBC>), gd, bd, lds, st, a, P); \
    else                                                                                 \
      hipLaunchKernelGGL((k_grid_step<V, make_plan(V).RL, T, RBC>), gd, bd, lds, st, a, P); \
    return true;
    MH_GRID_COMPILED(MH_CASE)
#undef MH_CASE
    default: return false;
  }
}

// Defined in gridstep_rb1.hip and gridstep_rb2.hip.
bool grid_launch_rb1(const GridArgs& a, const GPlan& P, dim3 gd, dim3 bd, size_t lds,
                     hipStream_t st);
bool grid_launch_rb2(const GridArgs& a, const GPlan& P, dim3 gd, dim3 bd, size_t lds,
                     hipStream_t st);
