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

// KPP's Rosenbrock family by IPAR[3] number: 1 Ros2, 2 Ros3, 3 Ros4,
// 4 Rodas3, 5 Rodas4 -- coefficients as published (Sandu, Verwer, Blom,
// Spee, Carmichael and Potra 1997, Atmos. Environ. 31; Hairer and Wanner
// 1996, RODAS).  A and C hold the strictly lower triangles packed by rows,
// entry (i, j), j < i, at i*(i-1)/2 + j; alpha: stage times, gamma: row
// sums of the gamma matrix (the non-autonomous terms).  The reference
// initialises only Ros3 (mgpu.cu:1008-1062, 3802-3811).
struct RosMethod {
  int S;
  double A[15], C[15], M[6], E[6];
  int newf[6];
  double gamma1, elo;
  double alpha[6], gamma[6];
};
constexpr double ROS2_G = 1.0 + 1.0 / 1.4142135623730951;  // 1 + 1/sqrt(2)
__device__ constexpr RosMethod kRosMethods[6] = {
    {},
    {2, {1.0 / ROS2_G}, {-2.0 / ROS2_G}, {3.0 / (2.0 * ROS2_G), 1.0 / (2.0 * ROS2_G)},
     {1.0 / (2.0 * ROS2_G), 1.0 / (2.0 * ROS2_G)}, {1, 1}, ROS2_G, 2.0, {0.0, 1.0},
     {ROS2_G, -ROS2_G}},
    {3, {1.0, 1.0, 0.0}, {ROS_C21, ROS_C31, ROS_C32}, {ROS_M1, ROS_M2, ROS_M3},
     {ROS_E1, ROS_E2, ROS_E3}, {1, 1, 0}, ROS_GAMMA1, ROS_ELO,
     {0.0, 0.43586652150845899941601945119356, 0.43586652150845899941601945119356},
     {0.43586652150845899941601945119356, 0.24291996454816804366592249683314,
      2.1851380027664058511513169485832}},
    {4,
     {2.0, 1.867943637803922, 0.2344449711399156, 1.867943637803922, 0.2344449711399156, 0.0},
     {-7.137615036412310, 2.580708087951457, 0.6515950076447975, -2.137148994382534,
      -0.3214669691237626, -0.6949742501781779},
     {2.255570073418735, 0.2870493262186792, 0.4353179431840180, 1.093502252409163},
     {-0.2815431932141155, -0.07276199124938920, -0.1082196201495311, -1.093502252409163},
     {1, 1, 1, 0},
     0.5728200000000000,
     4.0,
     {0.0, 1.145640000000000, 0.6552168638155900, 0.6552168638155900},
     {0.5728200000000000, -1.769193891319233, 0.7592633437920482, -0.1049021087100450}},
    {4, {0.0, 2.0, 0.0, 2.0, 0.0, 1.0}, {4.0, 1.0, -1.0, 1.0, -1.0, -(8.0 / 3.0)},
     {2.0, 0.0, 1.0, 1.0}, {0.0, 0.0, 0.0, 1.0}, {1, 0, 1, 1}, 0.5, 3.0, {0.0, 0.0, 1.0, 1.0},
     {0.5, 1.5, 0.0, 0.0}},
    {6,
     {1.544000000000000, 0.9466785280815826, 0.2557011698983284, 3.314825187068521,
      2.896124015972201, 0.9986419139977817, 1.221224509226641, 6.019134481288629,
      12.53708332932087, -0.6878860361058950, 1.221224509226641, 6.019134481288629,
      12.53708332932087, -0.6878860361058950, 1.0},
     {-5.668800000000000, -2.430093356833875, -0.2063599157091915, -0.1073529058151375,
      -9.594562251023355, -20.47028614809616, 7.496443313967647, -10.24680431464352,
      -33.99990352819905, 11.70890893206160, 8.083246795921522, -7.981132988064893,
      -31.52159432874371, 16.31930543123136, -6.058818238834054},
     {1.221224509226641, 6.019134481288629, 12.53708332932087, -0.6878860361058950, 1.0, 1.0},
     {0.0, 0.0, 0.0, 0.0, 0.0, 1.0},
     {1, 1, 1, 1, 1, 1},
     0.25,
     4.0,
     {0.0, 0.386, 0.210, 0.630, 1.0, 1.0},
     {0.25, -0.1043, 0.1035, -0.3620000000000023e-01, 0.0, 0.0}},
};

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
  int32_t method;        // KPP's IPAR[3] number: 1 Ros2, 2 Ros3, 3 Ros4, 4 Rodas3, 5 Rodas4
  int32_t sing_retries;  // step halvings after a singular matrix before -8
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

// The fused stages' decomposition and solves.  The L and U entries that
// the decomposition does not store are formed at their use (by the solves,
// and for U also by the rows that eliminate with them) from the same rate
// constants, concentrations and inverse pivots, by the same operations:
// the state and the rate constants they use (no photolysis rate, which the
// forcing moves between stages) are fixed through an attempt.  WIDE: the
// larger set (entries of up to three Jacobian terms instead of two).  Both
// pairs issue their loads 20 at a time.
template <bool WIDE, class KA, class YA, class GA>
__device__ __forceinline__ int stage_decomp(KA k, YA y, double h2o, double m, double ghinv,
                                            GA g) {
  if constexpr (WIDE) return gen::jac_decomp_rcw_pf20(k, y, h2o, m, ghinv, g);
  else return gen::jac_decomp_rc_pf20(k, y, h2o, m, ghinv, g);
}

template <bool WIDE, class KA, class YA, class GA>
__device__ __forceinline__ void stage_solve_reg(KA k, YA y, double h2o, double m, GA g,
                                                double (&x)[NVAR]) {
  if constexpr (WIDE) gen::solve_rcw_pf20_reg(k, y, h2o, m, g, x);
  else gen::solve_rc_pf20_reg(k, y, h2o, m, g, x);
}

template <int RF, class KA, class VA, class FA>
__device__ __forceinline__ void stage_fun(KA k, VA v, double h2o, double m, FA f) {
  if constexpr (is_ws_like<FA>::value) gen::fun_pf16(k, v, h2o, m, f);
  else gen::fun<RF>(k, v, h2o, m, f);
}

// The stage-2 state y + K1 read where the tendency uses it: entry i is
// RD(y, i) + RD(k1, i), the addition the stored stage state was formed
// with, so the state is neither written out nor read back.  In the
// load-scheduled code both operands are loaded with the batch and added at
// the use.
template <class YA, class KA>
struct SumView {
  YA y;
  KA k;
};
struct SumRead {
  double a, b;
};
template <int I, class YA, class KA>
__device__ __forceinline__ SumRead RDP(const SumView<YA, KA>& v) {
  return SumRead{V(RDP<I>(v.y)), RD(v.k, I)};
}
__device__ __forceinline__ double V(const SumRead& r) { return r.a + r.b; }
template <class YA, class KA>
__device__ __forceinline__ double RD(const SumView<YA, KA>& v, int i) {
  return RD(v.y, i) + RD(v.k, i);
}

// The stage-2 state y + a K1 of a method whose A(2,1) = a is not 1 (Ros2),
// read the same way: entry i is fma(a, RD(k1, i), RD(y, i)), the operation
// the general stage loop forms that state with.
template <class YA, class KA>
struct AxpyView {
  YA y;
  KA k;
  double a;
};
struct AxpyRead {
  double y, k, a;
};
template <int I, class YA, class KA>
__device__ __forceinline__ AxpyRead RDP(const AxpyView<YA, KA>& v) {
  return AxpyRead{V(RDP<I>(v.y)), RD(v.k, I), v.a};
}
__device__ __forceinline__ double V(const AxpyRead& r) { return fma(r.a, r.k, r.y); }
template <class YA, class KA>
__device__ __forceinline__ double RD(const AxpyView<YA, KA>& v, int i) {
  return fma(v.a, RD(v.k, i), RD(v.y, i));
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

// Photolysis forcing of a non-autonomous integration: J(t) = j0 + (t/dt)
// (j1 - j0) for this cell (j0, j1 at its column, stride ld); j1 null: the
// integration is autonomous.
struct Forcing {
  const double* j0;
  const double* j1;
  int64_t ld;
  double dt;
};

// The photolysis rate constants (a bare J each, mgpu.cu:763-783) at time t,
// J read as global memory.
template <class KA>
__device__ __forceinline__ void photo_at(KA k, const Forcing& fc, double t) {
  const double w = t / fc.dt;
  const auto j0 = gmem(fc.j0), j1 = gmem(fc.j1);
#pragma unroll
  for (int r = 0; r < gen::NPHOTO; r++) {
    const double a = j0[gen::PHOTO_J[r] * fc.ld], b = j1[gen::PHOTO_J[r] * fc.ld];
    WR(k, r, fma(w, b - a, a));
  }
}

// Step of the dF/dT difference quotient at time T: to the farther end of
// [0, dt].  The forcing is linear in t, so the quotient is its exact slope
// for any step; KPP's sqrt(roundoff) max(DeltaMin, |T|) would cancel to
// noise near T = 0, where every call's time axis starts.
__device__ __forceinline__ double forcing_delta(double T, double dt) {
  return T <= 0.5 * dt ? dt - T : -T;
}

// Stages of one attempt of KPP's generic Rosenbrock scheme on workspace
// views (the methods other than Ros3 and Ros2; forced Ros3 in a wave whose
// other cells are not all forced Ros3; forced Ros2, and autonomous Ros2 in a
// wave whose other cells are not all autonomous Ros2): stage
// s's vector lives in the workspace region k1.region + s*NVAR, dF/dT in
// region 6, the latest new F in f.  The reference's commented stage loop
// (mgpu.cu:3595-3625) in KPP's general form: F = f(T, y) for stage 0,
// f(T + alpha_s H, y + sum_j A(s,j) K_j) when NewF(s), else the previous F;
// K_s = G^-1 (F + sum_j C(s,j)/H K_j + H gamma_s dF/dT), the last term when
// forced (mgpu.cu:1566-1593, 3619-3622).  The rate constants k hold time T
// on entry.  As in the fused Ros3 stages, each right-hand side is formed in
// registers (the tendency writes it there) and solved there; a stage's K
// is stored once.  The caller hands one method at a time, the same for
// every lane that runs it, so the coefficients come through the scalar
// cache.  The loops over stages and over earlier stages run at run time
// (unrolling them for each stage count was slower: code size and spills);
// the loops over species are unrolled (constant element indices keep x and
// the hot set in registers and LDS).  Returns the sum of squared
// scaled errors (sum_j E_j K_j) and leaves Ynew = y + sum_j M_j K_j in f.
template <class KA, class YA, class GA, class VA>
__device__ __forceinline__ double generic_stages(const RosMethod& rm, const Forcing& fc, double T,
                                                 KA k, YA y, double h2o, double m, GA g, VA f,
                                                 VA k1, double H, double rtol, double atol) {
  const int S = rm.S;
  const bool forced = fc.j1 != nullptr;
  VA dfdt = k1;
  dfdt.region = k1.region + 6 * NVAR;
#pragma nounroll
  for (int st = 0; st < S; st++) {
    const int base = st * (st - 1) / 2;
    VA ks = k1;
    ks.region = k1.region + st * NVAR;
    double x[NVAR];
    if (st == 0 || rm.newf[st]) {
      // the stage's state: y for stage 0, y + sum_j A(s,j) K_j after,
      // accumulated in x and stored in the stage's region, so that one
      // instance of the tendency serves every stage
#pragma unroll
      for (int i = 0; i < NVAR; i++) x[i] = RD(y, i);
      for (int j = 0; j < st; j++) {
        VA kj = k1;
        kj.region = k1.region + j * NVAR;
        const double a = rm.A[base + j];
        xpass<16>([&x, a](int i, const double* v) { x[i] = fma(a, v[0], x[i]); }, kj);
      }
#pragma unroll
      for (int i = 0; i < NVAR; i++) WR(ks, i, x[i]);
      launder(ks);
      if (forced && st > 0) photo_at(k, fc, T + rm.alpha[st] * H);
      gen::fun_pf20_reg(k, ks, h2o, m, x);
#pragma unroll
      for (int i = 0; i < NVAR; i++) WR(f, i, x[i]);
      if (st == 0 && forced) {
        // dF/dT = (f(T + delta, y) - f(T, y)) / delta; x back to F0 after
        const double delta = forcing_delta(T, fc.dt);
        const double rd = 1.0 / delta;
        photo_at(k, fc, T + delta);
        gen::fun_pf20_reg(k, ks, h2o, m, x);
        VA f0 = f;
        launder(f0);
#pragma unroll
        for (int i = 0; i < NVAR; i++) {
          const double fi = RD(f0, i);
          WR(dfdt, i, fma(-1.0, fi, x[i]) * rd);
          x[i] = fi;
        }
      }
    } else {
      VA fv = f;
      launder(fv);
#pragma unroll
      for (int i = 0; i < NVAR; i++) x[i] = RD(fv, i);
    }
    // + sum_j C(s,j)/H K_j, then H gamma_s dF/dT
    for (int j = 0; j < st; j++) {
      VA kj = k1;
      kj.region = k1.region + j * NVAR;
      const double c = rm.C[base + j] / H;
      xpass<16>([&x, c](int i, const double* v) { x[i] = fma(v[0], c, x[i]); }, kj);
    }
    if (forced && rm.gamma[st] != 0.0) {
      const double hg = H * rm.gamma[st];
      VA d = dfdt;
      launder(d);
      xpass<16>([&x, hg](int i, const double* v) { x[i] = fma(hg, v[0], x[i]); }, d);
    }
    stage_solve_reg<false>(k, y, h2o, m, g, x);
#pragma unroll
    for (int i = 0; i < NVAR; i++) WR(ks, i, x[i]);
  }
  launder(k1);
  double err = 0.0;
  // combination and error norm, CH species at a time so that each stage's
  // loads for the chunk are issued together; per species the sums still run
  // over the stages in order, and err over the species in order
  constexpr int CH = 4;
#pragma unroll
  for (int c = 0; c < NVAR; c += CH) {
    double sm[CH], e[CH];
#pragma unroll
    for (int t = 0; t < CH; t++) {
      if (c + t < NVAR) {
        const double a0 = RD(k1, c + t);
        sm[t] = rm.M[0] * a0;
        e[t] = rm.E[0] * a0;
      }
    }
    for (int j = 1; j < S; j++) {
      VA kj = k1;
      kj.region = k1.region + j * NVAR;
      double aj[CH];
#pragma unroll
      for (int t = 0; t < CH; t++)
        if (c + t < NVAR) aj[t] = RD(kj, c + t);
#pragma unroll
      for (int t = 0; t < CH; t++) {
        if (c + t < NVAR) {
          sm[t] = fma(rm.M[j], aj[t], sm[t]);
          e[t] = fma(rm.E[j], aj[t], e[t]);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < CH; t++) {
      if (c + t < NVAR) {
        const int i = c + t;
        const double yi = RD(y, i);
        const double yn = yi + sm[t];
        const double ymax = dmax_ref(fabs(yi), fabs(yn));
        const double q = e[t] / fma(rtol, ymax, atol);
        err = fma(q, q, err);
        WR(f, i, yn);
      }
    }
  }
  return err;
}

// One attempt for a running cell whose attempt top has been applied.
//   y   state (updated on acceptance)     k   rate constants
//   g   659-entry LU workspace            f, k1, k2, k3: 59-entry vectors
// R: launder interval of the compiler-scheduled fused Jacobian/decomposition
// (rows) and, as 16 reactions per unit, of the tendency on private arrays;
// phases are laundered apart whenever the arrays are workspace views.
// WIDE: the Ros3 kernels' tuning of the fused stages (the wider set of
// formed LU entries, and F2 + C31/H K1 kept for the third stage instead of
// F2, which saves reloading K1 there); the family kernel, whose general
// stage loop shares the narrower set, was faster without either (measured
// at 1M cells; the one-shot grid was too until the load batches of 20).
// FAMILY: the kernel runs every method of KPP's Rosenbrock family, the
// non-autonomous forcing fc and KPP's step halving after a singular matrix
// (workspace views with seven stage regions); otherwise Ros3 alone,
// autonomous, -8 at once, whatever o and fc ask.
template <int R, bool FAMILY, bool WIDE, class KA, class YA, class GA, class VA>
__device__ __forceinline__ void ros3_attempt(CellState& s, const Ros3Opts& o, const Forcing& fc,
                                             KA k, YA y, double h2o, double m, GA g, VA f, VA k1,
                                             VA k2, VA k3) {
  static_assert(!FAMILY || is_ws_like<GA>::value, "the family needs workspace stage regions");
  static_assert(!(FAMILY && WIDE), "the general stage loop solves with the narrower set");
  constexpr int RF = R > 0 ? 16 * R : 0;
  double H = s.H;
  const bool forced = FAMILY && fc.j1 != nullptr;
  // forced Ros3 runs the fused stages with the forcing terms when every
  // active lane of the wave is a forced Ros3 cell: the condition is then
  // wave-uniform, so its branches move no lane mask (a mixed wave sends its
  // forced cells through the general stage loop, with identical results)
  bool fused_forced = false;
  if constexpr (FAMILY) fused_forced = __ballot(forced && o.method == 2) == __ballot(true);
  const bool ros3 = !FAMILY || (o.method == 2 && (!forced || fused_forced));
  // autonomous Ros2 runs the first two fused stages with its own
  // coefficients, then its own combination (its two stages have Ros3's
  // shape: a new F in each, the second at y + A21 K1), when every active
  // lane of the wave is an autonomous Ros2 cell (wave-uniform, as above; a
  // mixed wave sends its Ros2 cells through the general stage loop, with
  // identical results)
  bool two = false;
  if constexpr (FAMILY) two = __ballot(o.method == 1 && !forced) == __ballot(true);
  constexpr RosMethod r2 = kRosMethods[1];
  // the family reads its scalars from the table for every lane (Ros3's row
  // holds Ros3's values): a per-lane choice between a constant and a load
  // would be a branch whose merged value the compiler may spill under the
  // branch's lane mask only, leaving the other lanes' copy unset
  const RosMethod& rm = kRosMethods[FAMILY ? o.method : 2];
  const double elo = FAMILY ? rm.elo : ROS_ELO;
  const double gam = FAMILY ? rm.gamma1 : ROS_GAMMA1;
  const double ghinv = 1.0 / (H * gam);
  const double c21 = two ? r2.C[0] : ROS_C21;
  double hc21 = c21 / H, hc31 = ROS_C31 / H, hc32 = ROS_C32 / H;
  const double rtol = o.rtol, atol = o.atol;
  // On the Ros3 kernels the step controller's options are read at the top
  // with the others (read at their use they would wait a memory round trip
  // behind the attempt's stores); the family kernel reads them at their use
  // (its register budget is tighter: its attempts were 4-6% slower so)
  double facmin = 0.0, facsafe = 0.0, facmax = 0.0, hmin = 0.0, facrej = 0.0;
  if constexpr (!FAMILY) {
    facmin = o.facmin;
    facsafe = o.facsafe;
    facmax = o.facmax;
    hmin = o.hmin;
    facrej = o.facrej;
  }
  double err = 0.0;
  // the fused stages' right-hand side, solved in place (registers)
  double x[NVAR];
  if constexpr (is_ws_like<GA>::value) {
    // Fused stages: each right-hand side is formed in registers and solved
    // there; K1 and K2 are stored once, K3 never (the combination uses it
    // in place), and F2 (or, WIDE, F2 + C31/H K1) once for the third
    // stage.  The stage-2
    // state y + K1 is formed where the tendency reads it (never stored),
    // the third right-hand side takes K2 from the registers the second
    // solve leaves it in, and on the Ros3 kernels Ynew stays in registers
    // until the step is accepted.  Autonomous Ros2 runs the first two of
    // these stages with its own A21 and C21 and combines with K2 in
    // registers (K2 never stored).  The decomposition does not depend on
    // F0, so it comes first.
    // G = -J(y) + ghinv*I formed row by row inside its own LU decomposition
    if constexpr (FAMILY) {
      // the Jacobian at T; ros_PrepareMatrix (KPP): a singular matrix halves
      // H and retries, up to sing_retries times (the reference's active
      // code: none, mgpu.cu:3147-3163), then -8
      if (forced) photo_at(k, fc, s.T);
      for (int nsc = 0;; nsc++) {
        if (stage_decomp<WIDE>(k, y, h2o, m, 1.0 / (H * gam), g) == 0) break;
        if (nsc >= o.sing_retries) {
          s.H = H;
          s.status = ST_SINGULAR;
          return;
        }
        H = H * 0.5;
      }
      hc21 = c21 / H;
      hc31 = ROS_C31 / H;
      hc32 = ROS_C32 / H;
    } else {
      if (stage_decomp<WIDE>(k, y, h2o, m, ghinv, g) != 0) {
        s.status = ST_SINGULAR;
        return;
      }
    }
    if (!ros3 && !two) {
      // one method at a time, the same for every lane that runs it, so that
      // its coefficients are read through the scalar cache (a wave whose
      // cells share a method, the usual case, takes one round)
      bool pending = true;
      for (;;) {
        const unsigned long long want = __ballot(pending);
        if (want == 0ull) break;
        const int lead = __ffsll((long long)want) - 1;
        const int meth = __builtin_amdgcn_readfirstlane(__shfl(o.method, lead));
        // the table index through an opaque copy: inside the branch the
        // compiler would otherwise address the table with the per-lane
        // method it knows to be equal
        int tab = meth;
        asm volatile("" : "+s"(tab));
        if (pending && o.method == meth) {
          err = generic_stages(kRosMethods[tab], fc, s.T, k, y, h2o, m, g, f, k1, H, rtol, atol);
          pending = false;
        }
      }
    } else {
    // forced: dF/dT in the seventh stage region, the stage times and gamma
    // sums of Ros3, each term in the general stage loop's order
    VA dfdt = k1;
    dfdt.region = k1.region + 6 * NVAR;
    constexpr RosMethod r3 = kRosMethods[2];
    // K1 = G^-1 f(y)   (forced: + H gamma_1 dF/dT)
    if constexpr (FAMILY) {
      // forced: a second pass of the same tendency code at T + delta, F0
      // parked in f meanwhile
      const double delta = forcing_delta(s.T, fc.dt);
#pragma nounroll
      for (int pass = 0;; pass++) {
        gen::fun_pf20_reg(k, y, h2o, m, x);
        if (!fused_forced || pass > 0) break;
#pragma unroll
        for (int i = 0; i < NVAR; i++) WR(f, i, x[i]);
        photo_at(k, fc, s.T + delta);
      }
      if (fused_forced) {
        // dF/dT = (f(T + delta, y) - f(T, y)) / delta
        const double rd = 1.0 / delta;
        const double hg = H * r3.gamma[0];
        VA f0 = f;
        launder(f0);
#pragma unroll
        for (int i = 0; i < NVAR; i++) {
          const double fi = RD(f0, i);
          const double d = fma(-1.0, fi, x[i]) * rd;
          WR(dfdt, i, d);
          x[i] = fma(hg, d, fi);
        }
      }
    } else {
      gen::fun_pf20_reg(k, y, h2o, m, x);
    }
    stage_solve_reg<WIDE>(k, y, h2o, m, g, x);
#pragma unroll
    for (int i = 0; i < NVAR; i++) WR(k1, i, x[i]);
    // stage 2: F2 = f(y + K1); K2 = G^-1 (F2 + C21/H K1)
    //   (forced: F2 at T + alpha_2 H, + H gamma_2 dF/dT; Ros2: f(y + A21 K1))
    if constexpr (FAMILY) {
      if (fused_forced) photo_at(k, fc, s.T + r3.alpha[1] * H);
    }
    if constexpr (FAMILY) {
      // y + K1 as fma(1, K1, y), the same value, or Ros2's y + A21 K1
      gen::fun_pf20_reg(k, AxpyView<YA, VA>{y, k1, two ? r2.A[0] : 1.0}, h2o, m, x);
    } else {
      gen::fun_pf20_reg(k, SumView<YA, VA>{y, k1}, h2o, m, x);
    }
    if constexpr (WIDE) {
      // with K1 loaded: the second right-hand side F2 + C21/H K1, and the
      // inner term of the third's, F2 + C31/H K1, kept in f
      xpass<16>([&x, f, hc21, hc31](int i, const double* a) {
        const double f2 = x[i];
        WR(f, i, fma(a[0], hc31, f2));
        x[i] = fma(a[0], hc21, f2);
      }, k1);
    } else {
      if (!two) {
#pragma unroll
        for (int i = 0; i < NVAR; i++) WR(f, i, x[i]);
      }
      xpass<16>([&x, hc21](int i, const double* a) { x[i] = fma(a[0], hc21, x[i]); }, k1);
    }
    if constexpr (FAMILY) {
      if (fused_forced) {
        const double hg = H * r3.gamma[1];
        xpass<16>([&x, hg](int i, const double* a) { x[i] = fma(hg, a[0], x[i]); }, dfdt);
      }
    }
    stage_solve_reg<WIDE>(k, y, h2o, m, g, x);
    if (two) {
      // Ros2's combination and error norm with K2 in registers, in the
      // general stage loop's order; Ynew in f
      xpass<8>([&x, &err, f, rtol, atol](int i, const double* a) {
        const double a1 = a[0], yi = a[1];
        const double sm = fma(r2.M[1], x[i], r2.M[0] * a1);
        const double e = fma(r2.E[1], x[i], r2.E[0] * a1);
        const double yn = yi + sm;
        const double ymax = dmax_ref(fabs(yi), fabs(yn));
        const double q = e / fma(rtol, ymax, atol);
        err = fma(q, q, err);
        WR(f, i, yn);
      }, k1, y);
    } else {
#pragma unroll
      for (int i = 0; i < NVAR; i++) WR(k2, i, x[i]);
      // stage 3 reuses F2: K3 = G^-1 (F2 + C31/H K1 + C32/H K2), with K2
      // still in registers   (forced: + H gamma_3 dF/dT)
      if constexpr (WIDE) {
        xpass<16>([&x, hc32](int i, const double* a) { x[i] = fma(x[i], hc32, a[0]); }, f);
      } else {
        xpass<16>([&x, hc31, hc32](int i, const double* a) {
          x[i] = fma(x[i], hc32, fma(a[0], hc31, a[1]));
        }, k1, f);
      }
      if constexpr (FAMILY) {
        if (fused_forced) {
          const double hg = H * r3.gamma[2];
          xpass<16>([&x, hg](int i, const double* a) { x[i] = fma(hg, a[0], x[i]); }, dfdt);
        }
      }
      stage_solve_reg<WIDE>(k, y, h2o, m, g, x);
      // combination and WRMS error norm with K3 in registers; Ynew replaces
      // K3 in x on the Ros3 kernels (written to y on acceptance), in f on the
      // family's (whose other methods leave it there)
      xpass<8>([&x, &err, f, rtol, atol](int i, const double* a) {
        const double a1 = a[0], a2 = a[1], a3 = x[i], yi = a[2];
        const double sm = fma(ROS_M3, a3, fma(ROS_M2, a2, ROS_M1 * a1));
        const double yn = yi + sm;
        const double e = fma(ROS_E3, a3, fma(ROS_E2, a2, ROS_E1 * a1));
        const double ymax = dmax_ref(fabs(yi), fabs(yn));
        const double q = e / fma(rtol, ymax, atol);
        err = fma(q, q, err);
        if constexpr (FAMILY) WR(f, i, yn);
        else x[i] = yn;
      }, k1, k2, y);
    }
    }
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
  const double fsafe = dmax_ref(FAMILY ? o.facmin : facmin,
                                (FAMILY ? o.facsafe : facsafe) / pow(err, 1.0 / elo));
  const double fac = dmin_ref(FAMILY ? o.facmax : facmax, fsafe);
  double Hnew = H * fac;
  s.ns++;
  if (err <= 1.0 || H <= (FAMILY ? o.hmin : hmin)) {
    if constexpr (!FAMILY && is_ws_like<GA>::value) {
#pragma unroll
      for (int i = 0; i < NVAR; i++) WR(y, i, x[i]);
    } else {
      launder(f);
      launder(y);
      vmap<16>(y, [](int, const double* a) { return a[0]; }, f);
    }
    s.T = s.T + H;
    if (s.na == 0) s.hfirst = H * fsafe;  // uncapped proposal (learned-h0 target)
    if (s.rejLast) Hnew = dmin_ref(Hnew, H);
    s.rejLast = 0;
    s.rejMore = 0;
    s.H = Hnew;
    s.na++;
    s.newstep = 1;
  } else {
    if (s.rejMore) Hnew = H * (FAMILY ? o.facrej : facrej);
    s.rejMore = s.rejLast;
    s.rejLast = 1;
    s.H = Hnew;
    s.nr++;
  }
}

}  // namespace radm2
