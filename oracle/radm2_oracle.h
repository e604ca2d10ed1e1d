/*
 * CPU oracle for the RADM2 Rosenbrock (KPP Ros3) per-cell integrator.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP path and the
 * timed CPU baseline.  No product code links, loads or calls it.
 *
 * Layouts match the product C ABI: state arrays are structure-of-arrays with
 * the cell index fastest, x[s*ld + p].  Per-cell building blocks take one
 * cell's contiguous vectors.
 */
#ifndef RADM2_ORACLE_H
#define RADM2_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define O_NVAR 59
#define O_NREACT 156
#define O_LU_NNZ 659
#define O_NJ 21

/* sing_retries: step halvings allowed after a singular decomposition
 * (0 = the reference's active code, -8 at once; KPP: 5) */
typedef struct {
  double hmin, hmax, hstart, facmin, facmax, facrej, facsafe;
  int32_t max_steps, method, sing_retries;
} oracle_opts_in;

typedef struct {
  double hmin, hmax, hstart, facmin, facmax, facrej, facsafe, roundoff;
  int32_t max_steps, method, sing_retries; /* method: KPP number 1..5 */
} oracle_opts;

#define O_E_ARG (-101) /* negative sing_retries */

/* building blocks, one cell */
void oracle_rconst(double T, double M, const double *J21, double rc_n2o5, double *k);
double oracle_default_rc_n2o5(void);
void oracle_fun(const double *V, double h2o, double m, const double *k, double *vdot);
void oracle_jac(const double *V, double h2o, double m, const double *k, double *jvs);
void oracle_jac_prep(const double *V, double h2o, double m, const double *k, double ghinv,
                     double *G);
int oracle_decomp(double *G, int ref_division);
void oracle_solve(const double *LU, double *x, int ref_division);

/* option validation and defaults (KPP IERR codes -1..-5, 0 = ok) */
int oracle_resolve(double dt, double rtol, double atol, const oracle_opts_in *in,
                   oracle_opts *out);

/* one cell: returns status 1 ok, -6/-7/-8 */
int oracle_ros3_cell(double *y, const double *k, double h2o, double m, double dt, double rtol,
                     double atol, const oracle_opts *o, double h0, int32_t *nstep,
                     int32_t *nacc, int32_t *nrej, double *hexit);

/* many cells, SoA; returns the call-level code (0, -1..-5 or O_E_ARG).
 * jrate_end: NULL (autonomous) or the J rates at the end of the interval,
 * J(t) linear in between (non-autonomous). */
int oracle_integrate(int64_t n, int64_t ld, double *conc, const double *temp,
                     const double *cair, const double *h2o, const double *jrate,
                     const double *jrate_end, const double *rc_n2o5, double dt, double rtol,
                     double atol,
                     const oracle_opts_in *opts, const double *h0, int32_t *nstep,
                     int32_t *nacc, int32_t *nrej, int32_t *status, double *hexit,
                     int nthreads);

/* WRF-Chem fields -> solver inputs over a box (dev_convert, mgpu.cu:513-646);
 * grid = {nx, ny, nz, ox, oy, oz, bx, by, bz}, arrays [field][ny][nz][nx],
 * outputs as the product's radm2_wrf_convert. */
void oracle_wrf_convert(const int *grid, const double *phy, const double *moist, const double *ph,
                        const double *chem, const double *rad, int rc_fixed, int64_t ld,
                        double *conc, double *temp, double *cair, double *h2o, double *jrate,
                        double *rc_n2o5);

/* KPP's Rosenbrock family by IPAR[3] number (1 Ros2, 2 Ros3, 3 Ros4,
 * 4 Rodas3, 5 Rodas4): one cell with the generic stage loop.  j1 NULL:
 * autonomous (rate constants k throughout).  Otherwise non-autonomous: the
 * photolysis rate constants follow J(t) = j0 + (t/dt)(j1 - j0) (21 values
 * each, the reference's jv order), with KPP's finite-difference dF/dT. */
int oracle_ros_cell(int method, double *y, const double *k, const double *j0, const double *j1,
                    double h2o, double m, double dt, double rtol, double atol,
                    const oracle_opts *o, double h0, int32_t *nstep, int32_t *nacc,
                    int32_t *nrej, double *hexit);
int oracle_method_stages(int method);
/* a method's tables: A[15], C[15] (packed lower triangles), M[6], E[6],
 * alpha[6] (stage times), gamma[6] (row sums of its gamma matrix),
 * g_elo[2] = {gamma, ELO} */
void oracle_method_coeffs(int method, double *A, double *C, double *M, double *E,
                          double *alpha, double *gamma, double *g_elo);
/* acceptStep's step-size control (mgpu.cu:3238-3260) on the sum of squared
 * scaled errors: updates H, T and the two rejection flags; returns 1 when
 * the attempt is accepted, 0 when rejected. */
int oracle_step_control(double err_sum, double elo, const oracle_opts *o, double *H, double *T,
                        int32_t *rej_last, int32_t *rej_more);
/* one autonomous attempt of method 1..5 from y with step H: Ynew, the error
 * vector sum_j E_j K_j; returns the sum of squared scaled errors (-1:
 * singular matrix) */
double oracle_ros_attempt(int method, const double *y, const double *k, double h2o, double m,
                          double H, double rtol, double atol, double *ynew, double *yerr);
/* the same through the generic stage loop with a caller's coefficient table */
double oracle_ros_attempt_tab(int S, const double *A, const double *C, const double *M,
                              const double *E, const int32_t *newf, double gamma1,
                              const double *y, const double *k, double h2o, double m, double H,
                              double rtol, double atol, double *ynew, double *yerr);
/* photolysis rate constants at time t of [0, dt] into k */
void oracle_photo_at(double *k, const double *j0, const double *j1, double t, double dt);

#ifdef __cplusplus
}
#endif
#endif
