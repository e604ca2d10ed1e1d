/*
 * CPU oracle: restatement of the reference's per-cell RADM2 Rosenbrock
 * (KPP Ros3) integrator in plain C.
 *
 * TEST INFRASTRUCTURE ONLY.  It is the parity checker for the HIP path (its
 * functions are called only from tests/, the smoke check and the CPU-baseline
 * leg of bench.py) and is never linked into or called by the product.
 *
 * Reference semantics followed (file:line of the reference's mgpu.cu):
 *   rate-law helpers ARR2/TROE/TROEE/THERMAL_T2/k46  mgpu.cu:708-750
 *   rate constants                                   mgpu.cu:763-918
 *   tendency Fun (H2O2+OH loss omitted, :1275)        mgpu.cu:1103-1385
 *   Jacobian (keeps the H2O2+OH loss)                 mgpu.cu:1685-2651
 *   Ghimj = -Jac0 + 1/(H*gamma1) on the diagonal      mgpu.cu:3128-3149, 2685-2703
 *   row-wise no-pivot LU, zero-diagonal check         mgpu.cu:2726-2780
 *   forward/backward substitution                     mgpu.cu:2854-3096
 *   Ros3 coefficients                                 mgpu.cu:1008-1062
 *   stage formulas (the commented-out stage loop)     mgpu.cu:3595-3625
 *   combination, error norm, accept/reject            mgpu.cu:3223-3260
 *   time loop mask and H clamp                        mgpu.cu:1608-1613
 *   initial step, H <= 10*Roundoff -> 1e-6            mgpu.cu:3475-3477
 *   option defaults and validation                    mgpu.cu:3688-3811
 *   default call (Hstart = 0.01*dt, Ros3, autonomous) mgpu.cu:3850-3873
 *   unit roundoff (2^-52)                             mgpu.cu:949-973
 * Resolutions of the reference's defects: per-cell state everywhere; the
 * diagonal shift goes on the diagonal; FP64 M/E coefficients; the error
 * scale is atol + rtol*max(|Y|,|Ynew|); KPP's -6 (step count) and -7 (step
 * too small) checks are applied at the top of each time step.
 *
 * Arithmetic order contract shared with the HIP code generator: every
 * accumulation is fma(c, term, acc) from 0 in table order; monomials are
 * k times factors left to right; the LU stores inverse pivots unless the
 * reference's division form is asked for.
 */
#include "radm2_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- Ros3 (mgpu.cu:1008-1062) -------------------------------------------- */
static const double ROS_C21 = -1.0156171083877702091975600115545;
static const double ROS_C31 = 4.0759956452537699824805835358067;
static const double ROS_C32 = 9.2076794298330791242156818474003;
static const double ROS_M1 = 1.0;
static const double ROS_M2 = 6.1697947043828245592553615689730;
static const double ROS_M3 = -0.4277225654321857332623837380651;
static const double ROS_E1 = 0.5;
static const double ROS_E2 = -2.9079558716805469821718236208017;
static const double ROS_E3 = 0.2235406989781156962736090927619;
static const double ROS_GAMMA1 = 0.43586652150845899941601945119356;
static const double ROS_ELO = 3.0;
static const double DELTA_MIN = 1.0e-6;

#define O_MAX(a, b) (((a) >= (b)) ? (a) : (b))
#define O_MIN(b, c) (((b) < (c)) ? (b) : (c))

/* ---- rate-law helpers, as the reference writes them (mgpu.cu:708-750) ---- */
static double o_arr2(double A0, double B0, double T) { return A0 * exp(-B0 / T); }

static double o_troe(double k0_300K, double n, double kinf_300K, double m, double T, double M) {
  double zt = 300.0 / T;
  double k0 = k0_300K * pow(zt, n) * M;
  double kinf = kinf_300K * pow(zt, m);
  double kr = k0 / kinf;
  return k0 / (1.0 + kr) * pow(0.6, 1.0 / (1.0 + pow(log10(kr), 2)));
}

static double o_troee(double A, double B, double k0_300K, double n, double kinf_300K, double m,
                      double T, double M) {
  return A * exp(-B / T) * o_troe(k0_300K, n, kinf_300K, m, T, M);
}

static double o_thermal_t2(double c, double d, double T) { return pow(T, 2.0) * c * exp(-d / T); }

static double o_k46(double M, double T) {
  double k0 = 7.2e-15 * exp(785.0 / T);
  double k2 = 4.1e-16 * exp(1440.0 / T);
  double k3 = 1.9e-33 * exp(725.0 / T) * M;
  return k0 + k3 / (1 + k3 / k2);
}

/* mechanism tables and the literal rate laws (generated from the JSON) */
#include "generated/radm2_oracle_tables.h"

void oracle_rconst(double T, double M, const double *J21, double rc_n2o5, double *k) {
  o_rconst_literal(T, M, J21, rc_n2o5, k);
}

/* mgpu.cu:552 at relative humidity 0: 1/(3.6e4*exp(-(0/0.28)^2.8 + 300)) */
double oracle_default_rc_n2o5(void) { return 1.0 / (3.6e4 * exp(-pow(0.0 / 0.28, 2.8) + 300.0)); }

/* ---- tendency (mgpu.cu:1103-1385) ----------------------------------------- */
static inline double o_factor(int code, const double *V, double h2o, double m) {
  return code >= 100 ? (code == 100 ? h2o : m) : V[code];
}

void oracle_fun(const double *V, double h2o, double m, const double *k, double *vdot) {
  double A[O_NREACT];
  for (int r = 0; r < O_NREACT; r++) {
    double a = k[r];
    for (int q = 0; q < 4 && O_MONO[r][q] >= 0; q++) a = a * o_factor(O_MONO[r][q], V, h2o, m);
    A[r] = a;
  }
  for (int i = 0; i < O_NVAR; i++) {
    double acc = 0.0;
    for (int t = O_TEND_START[i]; t < O_TEND_START[i + 1]; t++)
      acc = fma(O_TEND_C[t], A[O_TEND_R[t]], acc);
    vdot[i] = acc;
  }
}

/* ---- Jacobian in LU layout (mgpu.cu:1685-2651) ---------------------------- */
static int o_lu_pos(int i, int j) {
  for (int kk = O_LU_CROW[i]; kk < O_LU_CROW[i + 1]; kk++)
    if (O_LU_ICOL[kk] == j) return kk;
  abort();
}

/* sign = -1 accumulates -J (prepared matrix), +1 accumulates J */
static void o_jac_acc(const double *V, double h2o, double m, const double *k, double sign,
                      double *G) {
  for (int p = 0; p < O_LU_NNZ; p++) G[p] = 0.0;
  for (int i = 0; i < O_NVAR; i++) {
    int nt = O_TEND_START[i + 1] - O_TEND_START[i];
    int nd = 0;
    for (int d = 0; d < O_NDROP; d++) nd += (O_DROP_I[d] == i);
    for (int t = 0; t < nt + nd; t++) {
      int r;
      double c;
      if (t < nt) {
        r = O_TEND_R[O_TEND_START[i] + t];
        c = O_TEND_C[O_TEND_START[i] + t];
      } else {
        int seen = -1, d;
        for (d = 0; d < O_NDROP; d++)
          if (O_DROP_I[d] == i && ++seen == t - nt) break;
        r = O_DROP_R[d];
        c = O_DROP_C[d];
      }
      for (int q = 0; q < 4 && O_MONO[r][q] >= 0; q++) {
        int code = O_MONO[r][q];
        if (code >= 100) continue;
        double term = k[r];
        for (int q2 = 0; q2 < 4 && O_MONO[r][q2] >= 0; q2++)
          if (q2 != q) term = term * o_factor(O_MONO[r][q2], V, h2o, m);
        int p = o_lu_pos(i, code);
        G[p] = fma(sign * c, term, G[p]);
      }
    }
  }
}

void oracle_jac(const double *V, double h2o, double m, const double *k, double *jvs) {
  o_jac_acc(V, h2o, m, k, 1.0, jvs);
}

void oracle_jac_prep(const double *V, double h2o, double m, const double *k, double ghinv,
                     double *G) {
  o_jac_acc(V, h2o, m, k, -1.0, G);
  for (int i = 0; i < O_NVAR; i++) G[O_LU_DIAG[i]] = G[O_LU_DIAG[i]] + ghinv;
}

/* ---- LU (mgpu.cu:2740-2779) and solve (mgpu.cu:2873-3090) ----------------- */
int oracle_decomp(double *G, int ref_division) {
  double W[O_NVAR];
  for (int k = 0; k < O_NVAR; k++) {
    if (G[O_LU_DIAG[k]] == 0.0) return k + 1;
    for (int kk = O_LU_CROW[k]; kk < O_LU_CROW[k + 1]; kk++) W[O_LU_ICOL[kk]] = G[kk];
    for (int kk = O_LU_CROW[k]; kk < O_LU_DIAG[k]; kk++) {
      int j = O_LU_ICOL[kk];
      double a;
      if (ref_division) { /* a = -W[j]/U_jj; W[j] = -a; W[i] += a*U_ji */
        a = -W[j] / G[O_LU_DIAG[j]];
        W[j] = -a;
      } else { /* inverse pivot on the diagonal */
        W[j] = W[j] * G[O_LU_DIAG[j]];
        a = -W[j];
      }
      for (int jj = O_LU_DIAG[j] + 1; jj < O_LU_CROW[j + 1]; jj++)
        W[O_LU_ICOL[jj]] = ref_division ? W[O_LU_ICOL[jj]] + a * G[jj]
                                        : fma(a, G[jj], W[O_LU_ICOL[jj]]);
    }
    for (int kk = O_LU_CROW[k]; kk < O_LU_CROW[k + 1]; kk++) G[kk] = W[O_LU_ICOL[kk]];
    if (!ref_division) G[O_LU_DIAG[k]] = 1.0 / W[k];
  }
  return 0;
}

void oracle_solve(const double *LU, double *x, int ref_division) {
  for (int i = 0; i < O_NVAR; i++)
    for (int kk = O_LU_CROW[i]; kk < O_LU_DIAG[i]; kk++)
      x[i] = ref_division ? x[i] - LU[kk] * x[O_LU_ICOL[kk]] : fma(-LU[kk], x[O_LU_ICOL[kk]], x[i]);
  for (int i = O_NVAR - 1; i >= 0; i--) {
    double acc = x[i];
    for (int kk = O_LU_DIAG[i] + 1; kk < O_LU_CROW[i + 1]; kk++)
      acc = ref_division ? acc - LU[kk] * x[O_LU_ICOL[kk]] : fma(-LU[kk], x[O_LU_ICOL[kk]], acc);
    x[i] = ref_division ? acc / LU[O_LU_DIAG[i]] : acc * LU[O_LU_DIAG[i]];
  }
}

/* ---- options (mgpu.cu:3688-3811, defaults of INTEGRATE mgpu.cu:3850-3873) - */
int oracle_resolve(double dt, double rtol, double atol, const oracle_opts_in *in,
                   oracle_opts *o) {
  oracle_opts_in d;
  if (!in) { /* INTEGRATE: all zero except Hstart = 0.01*dt, Ros3 */
    memset(&d, 0, sizeof d);
    d.hstart = 0.01 * dt;
    d.method = 2;
    in = &d;
  }
  double span = fabs(dt);
  o->roundoff = DBL_EPSILON;
  o->max_steps = in->max_steps == 0 ? 100000 : in->max_steps;
  if (o->max_steps < 0) return -1;
  /* method: 0 selects Ros3 (the reference's 0 fails its own range check) */
  int method = in->method == 0 ? 2 : in->method;
  if (method < 1 || method > 5) return -2;
  o->hmin = in->hmin;
  if (in->hmin < 0.0) return -3;
  o->hmax = in->hmax == 0.0 ? span : O_MIN(fabs(in->hmax), span);
  if (in->hmax < 0.0) return -3;
  o->hstart = in->hstart == 0.0 ? O_MAX(o->hmin, DELTA_MIN) : O_MIN(fabs(in->hstart), span);
  if (in->hstart < 0.0) return -3;
  o->facmin = in->facmin == 0.0 ? 0.2 : in->facmin;
  if (in->facmin < 0.0) return -4;
  o->facmax = in->facmax == 0.0 ? 6.0 : in->facmax;
  if (in->facmax < 0.0) return -4;
  o->facrej = in->facrej == 0.0 ? 0.1 : in->facrej;
  if (in->facrej < 0.0) return -4;
  o->facsafe = in->facsafe == 0.0 ? 0.9 : in->facsafe;
  if (in->facsafe < 0.0) return -4;
  if (atol <= 0.0 || rtol <= 10.0 * o->roundoff || rtol >= 1.0) return -5;
  /* the reference initialises only Ros3 (mgpu.cu:3802-3811); KPP's other
   * four methods are implemented here */
  o->method = method;
  if (in->sing_retries < 0) return O_E_ARG;
  o->sing_retries = in->sing_retries;
  return 0;
}

/* ---- step-size control (acceptStep, mgpu.cu:3238-3260) -------------------
 * err_sum: the sum over species of the squared scaled errors.  Err =
 * sqrt(err_sum/NVAR), Fac = min(FacMax, max(FacMin, FacSafe/Err^(1/ELO))),
 * Hnew = H*Fac.  Accepted when Err <= 1 or H <= Hmin: T += H, Hnew capped at
 * H after a rejection, both flags cleared.  Rejected: Hnew = H*FacRej after
 * two rejections in a row, RejectMoreH takes RejectLastH, RejectLastH set.
 * H takes Hnew either way.  Returns 1 when accepted, 0 when rejected. */
int oracle_step_control(double err_sum, double elo, const oracle_opts *o, double *H, double *T,
                        int32_t *rej_last, int32_t *rej_more) {
  const double err = sqrt(err_sum / (double)O_NVAR);
  const double fac = O_MIN(o->facmax, O_MAX(o->facmin, o->facsafe / pow(err, 1.0 / elo)));
  double Hnew = *H * fac;
  if (err <= 1.0 || *H <= o->hmin) {
    *T = *T + *H;
    if (*rej_last) Hnew = O_MIN(Hnew, *H);
    *rej_last = 0;
    *rej_more = 0;
    *H = Hnew;
    return 1;
  }
  if (*rej_more) Hnew = *H * o->facrej;
  *rej_more = *rej_last;
  *rej_last = 1;
  *H = Hnew;
  return 0;
}

/* ---- Ros3 stages of one attempt (mgpu.cu:3595-3625, combination 3223-3237) -
 * G: the decomposed -J + 1/(H gamma1) I; F0 = f(y).  Writes Ynew = y + sum
 * M_j K_j into yn and, when yerr is not NULL, sum E_j K_j into yerr;
 * returns the sum of squared scaled errors. */
static double o_ros3_stages(const double *y, const double *F0, const double *k, double h2o,
                            double m, double H, double rtol, double atol, const double *G,
                            double *yn, double *yerr) {
  double F2[O_NVAR], K1[O_NVAR], K2[O_NVAR], K3[O_NVAR];
  for (int i = 0; i < O_NVAR; i++) K1[i] = F0[i];
  oracle_solve(G, K1, 0);
  for (int i = 0; i < O_NVAR; i++) yn[i] = y[i] + K1[i];
  oracle_fun(yn, h2o, m, k, F2);
  const double hc21 = ROS_C21 / H, hc31 = ROS_C31 / H, hc32 = ROS_C32 / H;
  for (int i = 0; i < O_NVAR; i++) K2[i] = fma(K1[i], hc21, F2[i]);
  oracle_solve(G, K2, 0);
  for (int i = 0; i < O_NVAR; i++) K3[i] = fma(K2[i], hc32, fma(K1[i], hc31, F2[i]));
  oracle_solve(G, K3, 0);
  double err = 0.0;
  for (int i = 0; i < O_NVAR; i++) {
    double s = fma(ROS_M3, K3[i], fma(ROS_M2, K2[i], ROS_M1 * K1[i]));
    yn[i] = y[i] + s;
    double e = fma(ROS_E3, K3[i], fma(ROS_E2, K2[i], ROS_E1 * K1[i]));
    if (yerr) yerr[i] = e;
    double ymax = O_MAX(fabs(y[i]), fabs(yn[i]));
    double q = e / fma(rtol, ymax, atol);
    err = fma(q, q, err);
  }
  return err;
}

/* ---- one cell (CS3: mgpu.cu:3475-3625 with acceptStep 3223-3260) ---------- */
int oracle_ros3_cell(double *y, const double *k, double h2o, double m, double dt, double rtol,
                     double atol, const oracle_opts *o, double h0, int32_t *nstep,
                     int32_t *nacc, int32_t *nrej, double *hexit) {
  double G[O_LU_NNZ], F0[O_NVAR], yn[O_NVAR];
  const double Tend = dt, rnd = o->roundoff;
  double T = 0.0;
  double H = O_MIN(o->hstart, o->hmax);
  if (fabs(H) <= 10.0 * rnd) H = DELTA_MIN;
  if (h0 > 0.0) { /* per-cell user Hstart: mgpu.cu:3752, then 3475-3477 */
    H = O_MIN(O_MIN(h0, dt), o->hmax);
    if (fabs(H) <= 10.0 * rnd) H = DELTA_MIN;
  }
  int32_t rejLast = 0, rejMore = 0;
  int status = 1;
  int32_t ns = 0, na = 0, nr = 0;
  while ((T - Tend) + rnd <= 0.0) {
    int first = 1;
    oracle_fun(y, h2o, m, k, F0);
    for (;;) {
      /* KPP's -6/-7 checks, before every attempt (not only at the top of a
       * time step) so that endless rejections end; !(H > rnd) catches NaN */
      if (ns > o->max_steps) { status = -6; goto done; }
      if (T + 0.1 * H == T || !(H > rnd)) { status = -7; goto done; }
      if (first) { H = O_MIN(H, fabs(Tend - T)); first = 0; }
      const double ghinv = 1.0 / (H * ROS_GAMMA1);
      oracle_jac_prep(y, h2o, m, k, ghinv, G);
      if (oracle_decomp(G, 0) != 0) { status = -8; goto done; }
      const double err = o_ros3_stages(y, F0, k, h2o, m, H, rtol, atol, G, yn, NULL);
      ns++;
      if (oracle_step_control(err, ROS_ELO, o, &H, &T, &rejLast, &rejMore)) {
        for (int i = 0; i < O_NVAR; i++) y[i] = yn[i];
        na++;
        break;
      }
      nr++;
    }
  }
done:
  if (nstep) *nstep = ns;
  if (nacc) *nacc = na;
  if (nrej) *nrej = nr;
  if (hexit) *hexit = H;
  return status;
}

/* ---- KPP's Rosenbrock family, generic stage loop ----------------------------
 * Coefficients of KPP's methods by IPAR[3] number (1 Ros2, 2 Ros3, 3 Ros4,
 * 4 Rodas3, 5 Rodas4) as published: Sandu, Verwer, Blom, Spee, Carmichael
 * and Potra (1997), "Benchmarking stiff ODE solvers for atmospheric
 * chemistry problems II: Rosenbrock solvers", Atmos. Environ. 31, and
 * Hairer and Wanner (1996), Solving ODEs II (RODAS).  A and C are the
 * strictly lower triangles packed row by row: entry (i, j), j < i, at
 * i*(i-1)/2 + j.  The reference implements only Ros3 (mgpu.cu:1008-1062,
 * 3802-3811).  Stage scheme (the reference's commented stage loop,
 * mgpu.cu:3595-3625, in KPP's general form):
 *   F = f(y) for stage 0; F = f(y + sum_j A(s,j) K_j) when NewF(s), else
 *   the previous stage's F; K_s = G^-1 (F + sum_j C(s,j)/H K_j);
 *   Ynew = y + sum M_j K_j, error sum E_j K_j.
 * Non-autonomous (KPP, mgpu.cu:1566-1593 and 3619-3622): dF/dT by a
 * difference quotient (F(T + Delta) - F(T)) / Delta once per time step, a new
 * F of stage s evaluated at T + alpha_s H, and H gamma_s dF/dT added to stage
 * s's right-hand side.  alpha_i = sum_j alpha_ij and gamma_i = sum_j
 * gamma_ij of the published methods.  The forcing is linear in t, so any
 * Delta gives the exact slope; Delta reaches to the farther end of the
 * interval (KPP's sqrt(roundoff) max(DeltaMin, |T|) would cancel to noise
 * near T = 0, where every call's time axis starts).  The reference's own
 * time derivative evaluates Fun at T itself (its dF/dT is zero). */
typedef struct {
  int S;
  double A[15], C[15], M[6], E[6], alpha[6], gamma[6], gamma1, elo;
  int newf[6];
} o_method;

static o_method o_method_table(int method) {
  o_method r;
  memset(&r, 0, sizeof r);
  if (method == 1) { /* Ros2 */
    const double g = 1.0 + 1.0 / sqrt(2.0);
    r.S = 2;
    r.A[0] = 1.0 / g;
    r.C[0] = -2.0 / g;
    r.newf[0] = r.newf[1] = 1;
    r.M[0] = 3.0 / (2.0 * g);
    r.M[1] = 1.0 / (2.0 * g);
    r.E[0] = 1.0 / (2.0 * g);
    r.E[1] = 1.0 / (2.0 * g);
    r.alpha[1] = 1.0;
    r.gamma[0] = g;
    r.gamma[1] = -g;
    r.gamma1 = g;
    r.elo = 2.0;
  } else if (method == 2) { /* Ros3 */
    r.S = 3;
    r.A[0] = 1.0; r.A[1] = 1.0; r.A[2] = 0.0;
    r.C[0] = ROS_C21; r.C[1] = ROS_C31; r.C[2] = ROS_C32;
    r.newf[0] = 1; r.newf[1] = 1; r.newf[2] = 0;
    r.M[0] = ROS_M1; r.M[1] = ROS_M2; r.M[2] = ROS_M3;
    r.E[0] = ROS_E1; r.E[1] = ROS_E2; r.E[2] = ROS_E3;
    /* mgpu.cu:1054-1061 */
    r.alpha[1] = r.alpha[2] = 0.43586652150845899941601945119356;
    r.gamma[0] = 0.43586652150845899941601945119356;
    r.gamma[1] = 0.24291996454816804366592249683314;
    r.gamma[2] = 2.1851380027664058511513169485832;
    r.gamma1 = ROS_GAMMA1;
    r.elo = ROS_ELO;
  } else if (method == 3) { /* Ros4 */
    r.S = 4;
    r.A[0] = 2.0; r.A[1] = 1.867943637803922; r.A[2] = 0.2344449711399156;
    r.A[3] = r.A[1]; r.A[4] = r.A[2]; r.A[5] = 0.0;
    r.C[0] = -7.137615036412310; r.C[1] = 2.580708087951457; r.C[2] = 0.6515950076447975;
    r.C[3] = -2.137148994382534; r.C[4] = -0.3214669691237626; r.C[5] = -0.6949742501781779;
    r.newf[0] = r.newf[1] = r.newf[2] = 1; r.newf[3] = 0;
    r.M[0] = 2.255570073418735; r.M[1] = 0.2870493262186792;
    r.M[2] = 0.4353179431840180; r.M[3] = 1.093502252409163;
    r.E[0] = -0.2815431932141155; r.E[1] = -0.07276199124938920;
    r.E[2] = -0.1082196201495311; r.E[3] = -1.093502252409163;
    r.alpha[1] = 1.145640000000000;
    r.alpha[2] = r.alpha[3] = 0.6552168638155900;
    r.gamma[0] = 0.5728200000000000;
    r.gamma[1] = -1.769193891319233;
    r.gamma[2] = 0.7592633437920482;
    r.gamma[3] = -0.1049021087100450;
    r.gamma1 = 0.5728200000000000;
    r.elo = 4.0;
  } else if (method == 4) { /* Rodas3 */
    r.S = 4;
    r.A[0] = 0.0; r.A[1] = 2.0; r.A[2] = 0.0; r.A[3] = 2.0; r.A[4] = 0.0; r.A[5] = 1.0;
    r.C[0] = 4.0; r.C[1] = 1.0; r.C[2] = -1.0; r.C[3] = 1.0; r.C[4] = -1.0; r.C[5] = -(8.0 / 3.0);
    r.newf[0] = 1; r.newf[1] = 0; r.newf[2] = 1; r.newf[3] = 1;
    r.M[0] = 2.0; r.M[1] = 0.0; r.M[2] = 1.0; r.M[3] = 1.0;
    r.E[3] = 1.0;
    r.alpha[2] = r.alpha[3] = 1.0;
    r.gamma[0] = 0.5;
    r.gamma[1] = 1.5;
    r.gamma1 = 0.5;
    r.elo = 3.0;
  } else { /* Rodas4 */
    r.S = 6;
    const double a[15] = {1.544000000000000, 0.9466785280815826, 0.2557011698983284,
                          3.314825187068521, 2.896124015972201, 0.9986419139977817,
                          1.221224509226641, 6.019134481288629, 12.53708332932087,
                          -0.6878860361058950, 1.221224509226641, 6.019134481288629,
                          12.53708332932087, -0.6878860361058950, 1.0};
    const double c[15] = {-5.668800000000000, -2.430093356833875, -0.2063599157091915,
                          -0.1073529058151375, -9.594562251023355, -20.47028614809616,
                          7.496443313967647, -10.24680431464352, -33.99990352819905,
                          11.70890893206160, 8.083246795921522, -7.981132988064893,
                          -31.52159432874371, 16.31930543123136, -6.058818238834054};
    memcpy(r.A, a, sizeof a);
    memcpy(r.C, c, sizeof c);
    r.M[0] = a[6]; r.M[1] = a[7]; r.M[2] = a[8]; r.M[3] = a[9]; r.M[4] = 1.0; r.M[5] = 1.0;
    r.E[5] = 1.0;
    for (int i = 0; i < 6; i++) r.newf[i] = 1;
    r.alpha[1] = 0.386;
    r.alpha[2] = 0.210;
    r.alpha[3] = 0.630;
    r.alpha[4] = r.alpha[5] = 1.0;
    r.gamma[0] = 0.25;
    r.gamma[1] = -0.1043;
    r.gamma[2] = 0.1035;
    r.gamma[3] = -0.3620000000000023e-01;
    r.gamma1 = 0.25;
    r.elo = 4.0;
  }
  return r;
}

int oracle_method_stages(int method) { return o_method_table(method).S; }

void oracle_method_coeffs(int method, double *A, double *C, double *M, double *E,
                          double *alpha, double *gamma, double *g_elo) {
  const o_method r = o_method_table(method);
  memcpy(A, r.A, sizeof r.A);
  memcpy(C, r.C, sizeof r.C);
  memcpy(M, r.M, sizeof r.M);
  memcpy(E, r.E, sizeof r.E);
  memcpy(alpha, r.alpha, sizeof r.alpha);
  memcpy(gamma, r.gamma, sizeof r.gamma);
  g_elo[0] = r.gamma1;
  g_elo[1] = r.elo;
}

/* time step of the dF/dT difference quotient: to the farther end of [0, dt] */
static double o_forcing_delta(double T, double dt) { return T <= 0.5 * dt ? dt - T : -T; }

/* J(t) = j0 + (t/dt)(j1 - j0) for the photolysis reactions (rate constant
 * = the bare J, mgpu.cu:763-783) */
void oracle_photo_at(double *k, const double *j0, const double *j1, double t, double dt) {
  const double w = t / dt;
  for (int q = 0; q < O_NPHOTO; q++) {
    const int s = O_PHOTO_J[q];
    k[O_PHOTO_R[q]] = fma(w, j1[s] - j0[s], j0[s]);
  }
}

/* Stages of one attempt of the generic scheme with method table r: G is the
 * decomposed -J + 1/(H gamma1) I, F0 = f(T, y), kk the rate constants
 * (their photolysis entries are moved to T + alpha_s H when j1 is not NULL,
 * and dFdT enters each stage then).  Writes Ynew into yn and, when yerr is
 * not NULL, sum E_j K_j into yerr; returns the sum of squared scaled errors. */
static double o_ros_stages(const o_method *r, const double *y, const double *F0, double *kk,
                           const double *j0, const double *j1, const double *dFdT, double T,
                           double dt, double h2o, double m, double H, double rtol, double atol,
                           const double *G, double *yn, double *yerr) {
  double F[O_NVAR], K[6][O_NVAR];
  const int nonauto = j1 != NULL;
  for (int st = 0; st < r->S; st++) {
    const int base = st * (st - 1) / 2;
    if (st == 0) {
      memcpy(F, F0, sizeof F);
    } else if (r->newf[st]) {
      for (int i = 0; i < O_NVAR; i++) {
        double v = y[i];
        for (int j = 0; j < st; j++) v = fma(r->A[base + j], K[j][i], v);
        yn[i] = v;
      }
      if (nonauto) oracle_photo_at(kk, j0, j1, T + r->alpha[st] * H, dt);
      oracle_fun(yn, h2o, m, kk, F);
    }
    for (int i = 0; i < O_NVAR; i++) {
      double v = F[i];
      for (int j = 0; j < st; j++) v = fma(K[j][i], r->C[base + j] / H, v);
      K[st][i] = v;
    }
    if (nonauto && r->gamma[st] != 0.0) {
      const double hg = H * r->gamma[st];
      for (int i = 0; i < O_NVAR; i++) K[st][i] = fma(hg, dFdT[i], K[st][i]);
    }
    oracle_solve(G, K[st], 0);
  }
  double err = 0.0;
  for (int i = 0; i < O_NVAR; i++) {
    double sm = r->M[0] * K[0][i], e = r->E[0] * K[0][i];
    for (int j = 1; j < r->S; j++) {
      sm = fma(r->M[j], K[j][i], sm);
      e = fma(r->E[j], K[j][i], e);
    }
    yn[i] = y[i] + sm;
    if (yerr) yerr[i] = e;
    double ymax = O_MAX(fabs(y[i]), fabs(yn[i]));
    double q = e / fma(rtol, ymax, atol);
    err = fma(q, q, err);
  }
  return err;
}

int oracle_ros_cell(int method, double *y, const double *k, const double *j0, const double *j1,
                    double h2o, double m, double dt, double rtol, double atol,
                    const oracle_opts *o, double h0, int32_t *nstep, int32_t *nacc,
                    int32_t *nrej, double *hexit) {
  const o_method r = o_method_table(method);
  double G[O_LU_NNZ], F0[O_NVAR], yn[O_NVAR], dFdT[O_NVAR];
  double kk[O_NREACT];
  const int nonauto = j1 != NULL;
  memcpy(kk, k, sizeof kk);
  const double Tend = dt, rnd = o->roundoff;
  double T = 0.0;
  double H = O_MIN(o->hstart, o->hmax);
  if (fabs(H) <= 10.0 * rnd) H = DELTA_MIN;
  if (h0 > 0.0) {
    H = O_MIN(O_MIN(h0, dt), o->hmax);
    if (fabs(H) <= 10.0 * rnd) H = DELTA_MIN;
  }
  int32_t rejLast = 0, rejMore = 0;
  int status = 1;
  int32_t ns = 0, na = 0, nr = 0;
  while ((T - Tend) + rnd <= 0.0) {
    int first = 1;
    if (nonauto) oracle_photo_at(kk, j0, j1, T, dt);
    oracle_fun(y, h2o, m, kk, F0);
    if (nonauto) { /* ros_FunTimeDerivative with T perturbed */
      const double delta = o_forcing_delta(T, dt);
      oracle_photo_at(kk, j0, j1, T + delta, dt);
      oracle_fun(y, h2o, m, kk, dFdT);
      for (int i = 0; i < O_NVAR; i++) dFdT[i] = fma(-1.0, F0[i], dFdT[i]) * (1.0 / delta);
    }
    for (;;) {
      if (ns > o->max_steps) { status = -6; goto done; }
      if (T + 0.1 * H == T || !(H > rnd)) { status = -7; goto done; }
      if (first) { H = O_MIN(H, fabs(Tend - T)); first = 0; }
      if (nonauto) oracle_photo_at(kk, j0, j1, T, dt); /* the Jacobian at T */
      /* ros_PrepareMatrix (KPP): a singular matrix halves H and retries, up
       * to sing_retries times (the reference's active code: none,
       * mgpu.cu:3147-3163), then -8 */
      for (int nsc = 0;; nsc++) {
        const double ghinv = 1.0 / (H * r.gamma1);
        oracle_jac_prep(y, h2o, m, kk, ghinv, G);
        if (oracle_decomp(G, 0) == 0) break;
        if (nsc >= o->sing_retries) { status = -8; goto done; }
        H = H * 0.5;
      }
      const double err = o_ros_stages(&r, y, F0, kk, j0, j1, dFdT, T, dt, h2o, m, H, rtol, atol,
                                      G, yn, NULL);
      ns++;
      if (oracle_step_control(err, r.elo, o, &H, &T, &rejLast, &rejMore)) {
        for (int i = 0; i < O_NVAR; i++) y[i] = yn[i];
        na++;
        break;
      }
      nr++;
    }
  }
done:
  if (nstep) *nstep = ns;
  if (nacc) *nacc = na;
  if (nrej) *nrej = nr;
  if (hexit) *hexit = H;
  return status;
}

/* One autonomous attempt from y with step H (no step control): G = -J(y) +
 * 1/(H gamma1) I decomposed, F0 = f(y), then the method's stages.  method 2
 * runs the Ros3 stages oracle_ros3_cell uses, the others the generic loop.
 * Writes Ynew and sum_j E_j K_j; returns the sum of squared scaled errors,
 * or -1 when the matrix is singular. */
double oracle_ros_attempt(int method, const double *y, const double *k, double h2o, double m,
                          double H, double rtol, double atol, double *ynew, double *yerr) {
  const o_method r = o_method_table(method);
  double G[O_LU_NNZ], F0[O_NVAR], kk[O_NREACT];
  memcpy(kk, k, sizeof kk);
  oracle_jac_prep(y, h2o, m, kk, 1.0 / (H * r.gamma1), G);
  if (oracle_decomp(G, 0) != 0) return -1.0;
  oracle_fun(y, h2o, m, kk, F0);
  if (method == 2) return o_ros3_stages(y, F0, kk, h2o, m, H, rtol, atol, G, ynew, yerr);
  return o_ros_stages(&r, y, F0, kk, NULL, NULL, NULL, 0.0, 1.0, h2o, m, H, rtol, atol, G, ynew,
                      yerr);
}

/* oracle_ros_attempt through the generic loop with a caller's table (the
 * layout of oracle_method_coeffs: A[15], C[15], M[6], E[6], newf[6]). */
double oracle_ros_attempt_tab(int S, const double *A, const double *C, const double *M,
                              const double *E, const int32_t *newf, double gamma1,
                              const double *y, const double *k, double h2o, double m, double H,
                              double rtol, double atol, double *ynew, double *yerr) {
  o_method r;
  memset(&r, 0, sizeof r);
  r.S = S;
  memcpy(r.A, A, sizeof r.A);
  memcpy(r.C, C, sizeof r.C);
  memcpy(r.M, M, sizeof r.M);
  memcpy(r.E, E, sizeof r.E);
  for (int i = 0; i < 6; i++) r.newf[i] = newf[i];
  r.gamma1 = gamma1;
  double G[O_LU_NNZ], F0[O_NVAR], kk[O_NREACT];
  memcpy(kk, k, sizeof kk);
  oracle_jac_prep(y, h2o, m, kk, 1.0 / (H * gamma1), G);
  if (oracle_decomp(G, 0) != 0) return -1.0;
  oracle_fun(y, h2o, m, kk, F0);
  return o_ros_stages(&r, y, F0, kk, NULL, NULL, NULL, 0.0, 1.0, h2o, m, H, rtol, atol, G, ynew,
                      yerr);
}

int oracle_integrate(int64_t n, int64_t ld, double *conc, const double *temp,
                     const double *cair, const double *h2o, const double *jrate,
                     const double *jrate_end, const double *rc_n2o5, double dt, double rtol,
                     double atol,
                     const oracle_opts_in *opts, const double *h0, int32_t *nstep,
                     int32_t *nacc, int32_t *nrej, int32_t *status, double *hexit,
                     int nthreads) {
  oracle_opts o;
  int rc = oracle_resolve(dt, rtol, atol, opts, &o);
  if (rc != 0) return rc;
  const double rc_def = oracle_default_rc_n2o5();
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
#endif
  for (int64_t p = 0; p < n; p++) {
    double y[O_NVAR], k[O_NREACT], J[O_NJ], J1[O_NJ];
    for (int j = 0; j < O_NJ; j++) J[j] = jrate[j * ld + p];
    if (jrate_end)
      for (int j = 0; j < O_NJ; j++) J1[j] = jrate_end[j * ld + p];
    oracle_rconst(temp[p], cair[p], J, rc_n2o5 ? rc_n2o5[p] : rc_def, k);
    for (int i = 0; i < O_NVAR; i++) y[i] = conc[i * ld + p];
    int32_t a = 0, b = 0, c = 0;
    double hx = 0.0;
    const double h0p = h0 ? h0[p] : 0.0;
    int st = (o.method != 2 || o.sing_retries > 0 || jrate_end)
                 ? oracle_ros_cell(o.method, y, k, J, jrate_end ? J1 : NULL, h2o[p], cair[p], dt,
                                   rtol, atol, &o, h0p, &a, &b, &c, &hx)
                 : oracle_ros3_cell(y, k, h2o[p], cair[p], dt, rtol, atol, &o, h0p, &a, &b, &c,
                                    &hx);
    for (int i = 0; i < O_NVAR; i++) conc[i * ld + p] = y[i];
    if (nstep) nstep[p] = a;
    if (nacc) nacc[p] = b;
    if (nrej) nrej[p] = c;
    if (status) status[p] = st;
    if (hexit) hexit[p] = hx;
  }
  return 0;
}

/* ---- WRF-Chem input adapter: dev_convert (mgpu.cu:513-646) -------------- */
/* unit constants as mgpu.cu:30-44 defines them */
#define O_MWH2O (2 * O_MWH + O_MWO)
#define O_DENS2CON_A (1.0e-3 * (1.0 / O_MWAIR) * O_NAVGDRO)
#define O_DENS2CON_W (1.0e-3 * (1.0 / O_MWH2O) * O_NAVGDRO)

void oracle_wrf_convert(const int *g, const double *phy, const double *moist, const double *ph,
                        const double *chem, const double *rad, int rc_fixed, int64_t ld,
                        double *conc, double *temp, double *cair, double *h2o, double *jrate,
                        double *rc_n2o5) {
  const int nx = g[0], ny = g[1], nz = g[2], ox = g[3], oy = g[4], oz = g[5];
  const int bx = g[6], by = g[7], bz = g[8];
  const int64_t fs = (int64_t)nx * ny * nz;
  for (int j = 0; j < by; j++)
    for (int k = 0; k < bz; k++)
      for (int i = 0; i < bx; i++) {
        /* output p = j*gridDim.x*blockDim.x + k*blockDim.x + i (mgpu.cu:522),
         * input at the halo-shifted (j+oy, k+oz, i+ox) (mgpu.cu:532-535) */
        const int64_t p = ((int64_t)j * bz + k) * bx + i;
        const int64_t q = ((int64_t)(j + oy) * nz + (k + oz)) * nx + (i + ox);
        const double P = phy[q], T = phy[fs + q], rho = phy[2 * fs + q], qv = moist[fs + q];
        cair[p] = O_DENS2CON_A * rho;                        /* mgpu.cu:539 */
        h2o[p] = O_DENS2CON_W * qv * rho;                    /* mgpu.cu:542 */
        temp[p] = T;                                         /* mgpu.cu:545 */
        /* mgpu.cu:548-552 */
        const double es = 1000. * 0.6112 * exp(17.67 * (T - 273.15) / (T - 29.65));
        const double qvs = es / (P - es);
        double rh = qv / qvs;
        rh = O_MIN(O_MAX(rh, 0.0), 1.0);
        rc_n2o5[p] = rc_fixed ? 1.0 / (3.6e4 * exp(-pow(rh / 0.28, 2.8)) + 300.0)
                              : 1.0 / (3.6e4 * exp(-pow((rh / 0.28), 2.8) + 300.0));
        const double conv = 1.0E-6 * O_DENS2CON_A * rho;    /* mgpu.cu:555 */
        for (int s = 0; s < O_NJ; s++) jrate[s * ld + p] = ph[s * fs + q] / 60.0;  /* 557-584 */
        for (int s = 0; s < O_NVAR; s++) {                    /* mgpu.cu:587-645 */
          const double *a = O_WRF_SRC[s] ? rad : chem;
          conc[s * ld + p] = conv * O_MAX(a[O_WRF_FIELD[s] * fs + q], 0.0);
        }
      }
}
