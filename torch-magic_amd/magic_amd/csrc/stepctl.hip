// Step control on the device: the decisions the host takes between the
// radial and the lm phase of a CNAB2 step, so that a whole step (radial loop,
// dt decision, operator rebuild when dt changes, lm loop) is one stream of
// kernels and can be captured as one HIP graph.
//
//  * k_step_control (one workgroup): Courant limits per level from the
//    maxima the grid step accumulated, dtrkc = sqrt(delxr2 / max) and
//    dthkc = sqrt(delxh2 / max) with 1e10 where the maximum is 0
//    (courant.py:80-127); their minima; the hysteresis decision of
//    dt_courant (courant.py:130-155); the optional replayed step sizes; the
//    dt roll and rebuild decision of one_step (step_time.py:1574-1637,
//    rebuild on the first step or a new dt, as step_time.py:1628-1629); the
//    CNAB2 weights of set_weights (time_scheme.py:56-109); the first singular
//    system of the previous step's rebuild, if any.  Every expression
//    is evaluated in the order the host evaluates it, with no fused
//    multiply-add, so that host and device decisions and weights agree bit
//    for bit.
//  * k_fill: p[k*stride + i] = v, skipped when the rebuild flag is 0 (clears
//    operator blocks inside a captured step without a memset node).
#include "common.h"

#pragma clang fp contract(off)

__global__ __launch_bounds__(256) void k_step_control(MhStepCtlArgs a) {
  __shared__ double sr[256], sh[256];
  __shared__ int sbad[256];
  const int t = threadIdx.x;
  MhStepState* st = a.state;
  // pivot flags of the previous step's rebuild: first nonzero entry
  int bad = 0x7fffffff;
  if (a.info != nullptr && st->rebuild != 0 && st->sing_step == 0) {
    const int nflag = a.info_rows * a.info_cols;
    for (int i = t; i < nflag; i += 256)
      if (a.info[i] != 0 && i < bad) bad = i;
  }
  sbad[t] = bad;
  double mr = __builtin_inf(), mh = __builtin_inf();
  for (int r = t; r < a.N; r += 256) {
    const double vr = __longlong_as_double((long long)a.cfl[r]);
    const double vh = __longlong_as_double((long long)a.cfl[a.N + r]);
    const double cr = (vr > 0.0) ? __builtin_sqrt(a.delxr2[r] / vr) : 1e10;
    const double ch = (vh > 0.0) ? __builtin_sqrt(a.delxh2[r] / vh) : 1e10;
    mr = (cr < mr) ? cr : mr;
    mh = (ch < mh) ? ch : mh;
  }
  sr[t] = mr;
  sh[t] = mh;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      sr[t] = (sr[t + s] < sr[t]) ? sr[t + s] : sr[t];
      sh[t] = (sh[t + s] < sh[t]) ? sh[t + s] : sh[t];
      sbad[t] = (sbad[t + s] < sbad[t]) ? sbad[t + s] : sbad[t];
    }
    __syncthreads();
  }
  if (t != 0) return;
  if (sbad[0] != 0x7fffffff) {
    st->sing_step = st->n_step;
    st->sing_sys = sbad[0] / a.info_cols;
    st->sing_l = sbad[0] - st->sing_sys * a.info_cols;
    st->sing_info = a.info[sbad[0]];
  }
  const double dtrkc = sr[0], dthkc = sh[0];
  const int n = st->n_step + 1;
  const double dt = st->dt, dtMax = st->dtMax;
  // dt_courant: dtMin = min(dtrkc, dthkc); dt_2 = min(0.5 (1/dt_fac + 1) dtMin, dtMax)
  const double dt_fac = 2.0;
  const double dtMin = (dthkc < dtrkc) ? dthkc : dtrkc;
  const double x2 = 0.5 * (1.0 / dt_fac + 1.0) * dtMin;
  const double dt_2 = (dtMax < x2) ? dtMax : x2;
  bool new_dt;
  double dt_new;
  if (dt > dtMax) {
    new_dt = true;
    dt_new = dtMax;
  } else if (dt > dtMin) {
    new_dt = true;
    dt_new = dt_2;
  } else if (dt_fac * dt < dtMin && dt < dtMax) {
    new_dt = true;
    dt_new = dt_2;
  } else {
    new_dt = false;
    dt_new = dt;
  }
  if (a.dt_schedule != nullptr && n <= a.n_schedule) {
    dt_new = a.dt_schedule[n - 1];
    new_dt = dt_new != dt;
  }
  if (a.dt_log != nullptr && a.log_cap > 0) a.dt_log[(n - 1) % a.log_cap] = dt_new;
  const bool dirk = a.scheme == 1;
  const double dt1 = dt_new, dt2 = dirk ? dt_new : st->dt1;
  const double wl0 = dirk ? 0.5 * dt_new : a.alpha * dt_new;
  // step_time.py:1628-1629: rebuild on the first step or a new dt (and once
  // when no operators were built yet: wimp_lin0 is NaN)
  const bool rebuild = (n == 1) || new_dt || __builtin_isnan(st->wimp_lin0);
  if (rebuild) {
    st->wimp_lin0 = wl0;
    st->n_rebuild += 1;
  }
  if (dirk) {
    // BPR353 tableaux (time_scheme.py:186-200), scaled by dt (set_weights, :230-238)
    const double B[5][4] = {{0.0, 0.0, 0.0, 0.0},
                            {0.5, 0.5, 0.0, 0.0},
                            {5.0 / 18.0, -1.0 / 9.0, 0.5, 0.0},
                            {0.5, 0.0, 0.0, 0.5},
                            {0.25, 0.0, 0.75, -0.5}};
    const double E[5][3] = {{0.0, 0.0, 0.0},
                            {1.0, 0.0, 0.0},
                            {4.0 / 9.0, 2.0 / 9.0, 0.0},
                            {0.25, 0.0, 0.75},
                            {0.25, 0.0, 0.75}};
    for (int k = 1; k <= 4; ++k) {
      double* w = a.w_stage + 16 * (k - 1);
      for (int q = 0; q < 16; ++q) w[q] = 0.0;
      w[0] = 1.0;
      w[1] = B[k][0] * dt1;
      w[2] = E[k][0] * dt1;
      w[3] = E[k][1] * dt1;
      w[4] = 0.5 * dt1 * a.BuoFac;
      w[8] = B[k][1] * dt1;
      w[9] = E[k][2] * dt1;
      w[10] = B[k][2] * dt1;
      w[11] = B[k][3] * dt1;
    }
  } else {
    a.w[0] = 1.0;
    a.w[1] = (1.0 - a.alpha) * dt1;
    a.w[2] = (1.0 + 0.5 * dt1 / dt2) * dt1;
    a.w[3] = -0.5 * dt1 * dt1 / dt2;
    a.w[4] = a.alpha * dt1 * a.BuoFac;
  }
  st->dt1 = dt1;
  st->dt2 = dt2;
  st->time = st->time + dt_new;
  st->dt = dt_new;
  st->n_step = n;
  st->rebuild = rebuild ? 1 : 0;
  st->new_dt = new_dt ? 1 : 0;
}

extern "C" int mh_step_control(const MhStepCtlArgs* args, void* stream) {
  const MhStepCtlArgs& a = *args;
  if (a.N < 1 || a.cfl == nullptr || a.state == nullptr || a.w == nullptr || a.delxr2 == nullptr ||
      a.delxh2 == nullptr || (a.dt_schedule != nullptr && a.n_schedule < 0) ||
      (a.dt_log != nullptr && a.log_cap < 1) ||
      (a.info != nullptr && (a.info_rows < 1 || a.info_cols < 1)) ||
      (a.scheme != 0 && a.scheme != 1) || (a.scheme == 1 && a.w_stage == nullptr))
    return MH_EINVAL;
  hipLaunchKernelGGL(k_step_control, dim3(1), dim3(256), 0, mh_stream(stream), a);
  return mh_launch_status();
}

__global__ void k_fill(double* p, long n, long stride, double v, const int* when) {
  if (when && *when == 0) return;
  double* q = p + (long)blockIdx.y * stride;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    q[i] = v;
}

extern "C" int mh_fill_f64(double* p, long n, int count, long stride, double v, const int* when,
                           void* stream) {
  if (n < 0 || count < 0 || (n > 0 && count > 0 && p == nullptr) || (count > 1 && stride < n))
    return MH_EINVAL;
  if (n == 0 || count == 0) return 0;
  const long blocks = (n + 255) / 256;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)(blocks < 1024 ? blocks : 1024), (unsigned)count),
                     dim3(256), 0, mh_stream(stream), p, n, stride, v, when);
  return mh_launch_status();
}
