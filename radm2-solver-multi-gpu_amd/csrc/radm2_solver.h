// Per-cell FP64 Rosenbrock (KPP Ros3) step for RADM2, one cell per lane.
//
// Device-side building blocks shared by the one-shot grid kernel and the
// persistent work-stealing kernel.  The mechanism numerics (rate constants,
// tendency, prepared Jacobian, unrolled LU and solve) come from the
// generated header; this file holds the Ros3 stage combination, the error
// norm and the accept/reject state machine.
//
// Reference semantics (mgpu.cu of the reference):
//   Ros3 coefficients                         mgpu.cu:1008-1062
//   stage formulas (commented stage loop)     mgpu.cu:3595-3625, WAXPY_HC 1540-1552
//   prepared matrix -Jac0 + 1/(H*gamma1)      mgpu.cu:3128-3149
//   combination / error norm / step control   mgpu.cu:3223-3260
//   time-loop mask and H clamp                mgpu.cu:1608-1613
//   initial step                              mgpu.cu:3475-3477
// The arithmetic order matches the CPU oracle's (explicit fma, no
// contraction), so identical rate constants give identical bits.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "radm2_ws.h"
#include "generated/radm2_gen.h"
#include "radm2_hot.h"

namespace radm2 {

using gen::LU_NNZ;
using gen::NJ;
using gen::NREACT;
using gen::NVAR;

// Ros3 (Sandu et al. 1997 as coded in KPP)
constexpr double ROS_C21 = -1.0156171083877702091975600115545;
constexpr double ROS_C31 = 4.0759956452537699824805835358067;
constexpr double ROS_C32 = 9.2076794298330791242156818474003;
constexpr double ROS_M1 = 1.0;
constexpr double ROS_M2 = 6.1697947043828245592553615689730;
constexpr double ROS_M3 = -0.4277225654321857332623837380651;
constexpr double ROS_E1 = 0.5;
constexpr double ROS_E2 = -2.9079558716805469821718236208017;
constexpr double ROS_E3 = 0.2235406989781156962736090927619;
constexpr double ROS_GAMMA1 = 0.43586652150845899941601945119356;
constexpr double ROS_ELO = 3.0;
constexpr double DELTA_MIN = 1.0e-6;

// Per-cell status codes (KPP IERR)
constexpr int ST_RUNNING = 0;
constexpr int ST_OK = 1;
constexpr int ST_TOO_MANY_STEPS = -6;
constexpr int ST_STEP_TOO_SMALL = -7;
constexpr int ST_SINGULAR = -8;

// Resolved integration options for one domain (validated on the host).
struct Ros3Opts {
  double dt, rtol, atol;
  double hmin, hmax, hstart, facmin, facmax, facrej, facsafe, roundoff;
  int32_t max_steps;
  int32_t pad;
};

__device__ __forceinline__ double dmax_ref(double a, double b) { return a >= b ? a : b; }
__device__ __forceinline__ double dmin_ref(double b, double c) { return b < c ? b : c; }

// Scalar state of one cell between attempts.
struct CellState {
  double T, H;
  double hfirst;    // step-size proposal after the first accepted step
  int32_t ns, na, nr;
  int32_t status;   // ST_RUNNING while integrating
  int32_t rejLast, rejMore;
  int32_t newstep;  // 1 when the next attempt starts a new time step
};

__device__ __forceinline__ void cell_begin(CellState& s, const Ros3Opts& o, double h0) {
  s.T = 0.0;
  s.hfirst = 0.0;
  double H = dmin_ref(o.hstart, o.hmax);
  if (fabs(H) <= 10.0 * o.roundoff) H = DELTA_MIN;
  // a per-cell h0 is a user-selected Hstart for that cell: Hstart =
  // min(h0, dt) (mgpu.cu:3752), H = min(Hstart, Hmax) (mgpu.cu:3475-3477)
  if (h0 > 0.0) {
    H = dmin_ref(dmin_ref(h0, o.dt), o.hmax);
    if (fabs(H) <= 10.0 * o.roundoff) H = DELTA_MIN;
  }
  s.H = H;
  s.ns = s.na = s.nr = 0;
  s.status = ST_RUNNING;
  s.rejLast = s.rejMore = 0;
  s.newstep = 1;
}

// Top of an attempt.  At the start of a time step: the end-of-interval test
// (mgpu.cu:1608-1609).  Before every attempt: KPP's -6 (step count) and -7
// (step too small) checks; applying them to rejected attempts too bounds a
// cell whose attempts are never accepted, and !(H > roundoff) also catches a
// NaN step.  At the start of a time step H is clamped to the interval end
// (mgpu.cu:1612).
__device__ __forceinline__ void cell_attempt_top(CellState& s, const Ros3Opts& o) {
  const double Tend = o.dt;
  if (s.newstep && !((s.T - Tend) + o.roundoff <= 0.0)) {
    s.status = ST_OK;
    return;
  }
  if (s.ns > o.max_steps) {
    s.status = ST_TOO_MANY_STEPS;
    return;
  }
  if (s.T + 0.1 * s.H == s.T || !(s.H > o.roundoff)) {
    s.status = ST_STEP_TOO_SMALL;
    return;
  }
  if (s.newstep) {
    s.H = dmin_ref(s.H, fabs(Tend - s.T));
    s.newstep = 0;
  }
}

// On workspace views the solves, the fused Jacobian/decomposition and the
// tendency use the generated variants whose loads are issued 16 at a time
// ahead of their use (the memory round trips of a lane overlap instead of
// following one another); on private arrays the compiler keeps its own
// schedule.  Both compute the same operations in the same order.
template <class GA, class XA>
__device__ __forceinline__ void stage_solve(GA g, XA x) {
  if constexpr (is_ws_like<GA>::value) gen::solve_pf16(g, x);
  else gen::solve(g, x);
}

template <int R, class KA, class YA, class GA>
__device__ __forceinline__ int stage_jac_decomp(KA k, YA y, double h2o, double m, double ghinv,
                                                GA g) {
  if constexpr (is_ws_like<GA>::value) return gen::jac_decomp_pf16(k, y, h2o, m, ghinv, g);
  else return gen::jac_decomp<R>(k, y, h2o, m, ghinv, g);
}

template <int RF, class KA, class VA, class FA>
__device__ __forceinline__ void stage_fun(KA k, VA v, double h2o, double m, FA f) {
  if constexpr (is_ws_like<FA>::value) gen::fun_pf16(k, v, h2o, m, f);
  else gen::fun<RF>(k, v, h2o, m, f);
}

// Element-wise pass out[i] = fn(i, {in0[i], in1[i], ..}) over the 59
// entries, CH at a time: all inputs of a chunk are loaded before any of its
// outputs is stored, and the next chunk's inputs before this chunk's stores
// (an output never aliases an input), so on workspace views a pass costs
// about 59/CH memory round trips instead of one per entry.  fn runs for
// i = 0, 1, .., 58 in that order.
template <int CH, class A>
__device__ __forceinline__ void vload(double (&dst)[CH], const A& a, int c) {
#pragma unroll
  for (int j = 0; j < CH; j++)
    if (c * CH + j < NVAR) dst[j] = RD(a, c * CH + j);
}

template <int CH, class OA, class Fn, class... IA>
__device__ __forceinline__ void vmap(OA out, Fn fn, const IA&... in) {
  constexpr int N = sizeof...(IA);
  constexpr int NCH = (NVAR + CH - 1) / CH;
  double buf[2][N][CH];
  {
    int t = 0;
    (vload<CH>(buf[0][t++], in, 0), ...);
  }
#pragma unroll
  for (int c = 0; c < NCH; c++) {
    const int cur = c & 1;
    if (c + 1 < NCH) {
      int t = 0;
      (vload<CH>(buf[cur ^ 1][t++], in, c + 1), ...);
    }
#pragma unroll
    for (int j = 0; j < CH; j++) {
      const int i = c * CH + j;
      if (i < NVAR) {
        double v[N];
#pragma unroll
        for (int t = 0; t < N; t++) v[t] = buf[cur][t][j];
        WR(out, i, fn(i, v));
      }
    }
  }
}

// One Ros3 attempt for a running cell whose attempt top has been applied.
//   y   state (updated on acceptance)     k   rate constants
//   g   659-entry LU workspace            f, k1, k2, k3: 59-entry vectors
// R: launder interval of the compiler-scheduled fused Jacobian/decomposition
// (rows) and, as 16 reactions per unit, of the tendency on private arrays;
// phases are laundered apart whenever the arrays are workspace views.
// fn(i, {in0[i], in1[i], ..}) for i = 0, 1, .., 58 in order, the inputs
// streamed CH at a time with the next chunk's loads issued before this
// chunk's work (fenced, so the compiler keeps them apart); fn may store to
// views that none of the inputs alias.
template <int CH, class Fn, class... IA>
__device__ __forceinline__ void xpass(Fn fn, const IA&... in) {
  constexpr int N = sizeof...(IA);
  constexpr int NCH = (NVAR + CH - 1) / CH;
  double buf[2][N][CH];
  {
    int t = 0;
    (vload<CH>(buf[0][t++], in, 0), ...);
  }
#pragma unroll
  for (int c = 0; c < NCH; c++) {
    const int cur = c & 1;
    if (c + 1 < NCH) {
      int t = 0;
      (vload<CH>(buf[cur ^ 1][t++], in, c + 1), ...);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < CH; j++) {
      const int i = c * CH + j;
      if (i < NVAR) {
        double v[N];
#pragma unroll
        for (int t = 0; t < N; t++) v[t] = buf[cur][t][j];
        fn(i, v);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <int R, class KA, class YA, class GA, class VA>
__device__ __forceinline__ void ros3_attempt(CellState& s, const Ros3Opts& o, KA k, YA y,
                                             double h2o, double m, GA g, VA f, VA k1, VA k2,
                                             VA k3) {
  constexpr int RF = R > 0 ? 16 * R : 0;
  const double H = s.H;
  const double ghinv = 1.0 / (H * ROS_GAMMA1);
  const double hc21 = ROS_C21 / H, hc31 = ROS_C31 / H, hc32 = ROS_C32 / H;
  const double rtol = o.rtol, atol = o.atol;
  double err = 0.0;
  if constexpr (is_ws_like<GA>::value) {
    // Fused stages: each right-hand side is formed in registers and solved
    // there; K1 and K2 are stored once, K3 never (the combination uses it
    // in place), and F2 once for the third stage.  The decomposition does
    // not depend on F0, so it comes first.
    // G = -J(y) + ghinv*I formed row by row inside its own LU decomposition
    if (gen::jac_decomp_pf16(k, y, h2o, m, ghinv, g) != 0) {
      s.status = ST_SINGULAR;
      return;
    }
    double x[NVAR];
    // K1 = G^-1 f(y)
    gen::fun_pf16_reg(k, y, h2o, m, x);
    gen::solve_pf16_reg(g, x);
#pragma unroll
    for (int i = 0; i < NVAR; i++) WR(k1, i, x[i]);
    // stage 2: F2 = f(y + K1); K2 = G^-1 (F2 + C21/H K1)
    vmap<16>(k2, [](int, const double* a) { return a[0] + a[1]; }, y, k1);
    gen::fun_pf16_reg(k, k2, h2o, m, x);
#pragma unroll
    for (int i = 0; i < NVAR; i++) WR(f, i, x[i]);
    xpass<16>([&x, hc21](int i, const double* a) { x[i] = fma(a[0], hc21, x[i]); }, k1);
    gen::solve_pf16_reg(g, x);
#pragma unroll
    for (int i = 0; i < NVAR; i++) WR(k2, i, x[i]);
    // stage 3 reuses F2: K3 = G^-1 (F2 + C31/H K1 + C32/H K2)
    xpass<16>([&x, hc31, hc32](int i, const double* a) {
      x[i] = fma(a[0], hc32, fma(a[1], hc31, a[2]));
    }, k2, k1, f);
    gen::solve_pf16_reg(g, x);
    // combination and WRMS error norm with K3 in registers; Ynew kept in f
    xpass<8>([&x, &err, f, rtol, atol](int i, const double* a) {
      const double a1 = a[0], a2 = a[1], a3 = x[i], yi = a[2];
      const double sm = fma(ROS_M3, a3, fma(ROS_M2, a2, ROS_M1 * a1));
      const double yn = yi + sm;
      const double e = fma(ROS_E3, a3, fma(ROS_E2, a2, ROS_E1 * a1));
      const double ymax = dmax_ref(fabs(yi), fabs(yn));
      const double q = e / fma(rtol, ymax, atol);
      err = fma(q, q, err);
      WR(f, i, yn);
    }, k1, k2, y);
  } else {
    // F0 = f(y); K1 = G^-1 F0
    stage_fun<RF>(k, y, h2o, m, k1);
    // G = -J(y) + ghinv*I formed row by row inside its own LU decomposition
    if (stage_jac_decomp<R>(k, y, h2o, m, ghinv, g) != 0) {
      s.status = ST_SINGULAR;
      return;
    }
    stage_solve(g, k1);
    // stage 2: F2 = f(y + K1); K2 = G^-1 (F2 + C21/H K1)
    vmap<16>(k2, [](int, const double* a) { return a[0] + a[1]; }, y, k1);
    stage_fun<RF>(k, k2, h2o, m, f);
    vmap<16>(k2, [hc21](int, const double* a) { return fma(a[0], hc21, a[1]); }, k1, f);
    stage_solve(g, k2);
    // stage 3 reuses F2: K3 = G^-1 (F2 + C31/H K1 + C32/H K2)
    vmap<16>(k3, [hc31, hc32](int, const double* a) {
      return fma(a[0], hc32, fma(a[1], hc31, a[2]));
    }, k2, k1, f);
    stage_solve(g, k3);
    // combination and WRMS error norm; Ynew kept in f until the decision
    vmap<8>(f, [&err, rtol, atol](int, const double* a) {
      const double a1 = a[0], a2 = a[1], a3 = a[2], yi = a[3];
      const double sm = fma(ROS_M3, a3, fma(ROS_M2, a2, ROS_M1 * a1));
      const double yn = yi + sm;
      const double e = fma(ROS_E3, a3, fma(ROS_E2, a2, ROS_E1 * a1));
      const double ymax = dmax_ref(fabs(yi), fabs(yn));
      const double q = e / fma(rtol, ymax, atol);
      err = fma(q, q, err);
      return yn;
    }, k1, k2, k3, y);
  }
  err = sqrt(err / (double)NVAR);
  const double fsafe = dmax_ref(o.facmin, o.facsafe / pow(err, 1.0 / ROS_ELO));
  const double fac = dmin_ref(o.facmax, fsafe);
  double Hnew = H * fac;
  s.ns++;
  if (err <= 1.0 || H <= o.hmin) {
    launder(f);
    launder(y);
    vmap<16>(y, [](int, const double* a) { return a[0]; }, f);
    s.T = s.T + H;
    if (s.na == 0) s.hfirst = H * fsafe;  // uncapped proposal (learned-h0 target)
    if (s.rejLast) Hnew = dmin_ref(Hnew, H);
    s.rejLast = 0;
    s.rejMore = 0;
    s.H = Hnew;
    s.na++;
    s.newstep = 1;
  } else {
    if (s.rejMore) Hnew = H * o.facrej;
    s.rejMore = s.rejLast;
    s.rejLast = 1;
    s.H = Hnew;
    s.nr++;
  }
}

}  // namespace radm2
