/*
 * radm2.h -- C ABI of the MI355X-native RADM2 Rosenbrock integrator.
 *
 * One call integrates many independent grid cells of the RADM2 mechanism
 * (59 variable species, 2 fixed, 156 reactions) over [0, dt] with the KPP
 * Ros3 method, one cell per GPU lane, entirely on the device.
 *
 * The reference (Rizkyanno/RADM2-Solver-Multi-GPU, a single CUDA file) has no
 * library boundary: its integrator is the KPP C driver
 *     int Rosenbrock(double Tstart, double Tend, double *AbsTol,
 *                    double *RelTol, double RPAR[20], int IPAR[20]);
 * (mgpu.cu:3664-3848), reached from INTEGRATE() (mgpu.cu:3850-3889), with the
 * cell state passed through globals (DataDevice[0].d_VAR / d_FIX / d_RCONST,
 * mgpu.cu:263-292) filled by dev_convert / dev_Update_RCONST
 * (mgpu.cu:513-646, 752-919).  Each entry point below names the reference
 * interface it replaces.
 *
 * Conventions
 *  - Arrays are device pointers, structure-of-arrays with the cell index
 *    fastest: element (s, p) of a [S][ld] array is x[s*ld + p], p < ncells.
 *  - Species order is the reference's ind_* order (mgpu.cu:125-183); the 21
 *    photolysis rates are in the reference's jv order (mgpu.cu:763-783), s^-1.
 *  - Calls are asynchronous on the context's stream; radm2_sync() waits.
 *  - Call-level return: 0 on success; a KPP option error -1..-5 (raised
 *    before any launch, as mgpu.cu:3715-3811 does); or a RADM2_E_* code.
 *    Per-cell failures (-6 too many steps, -7 step too small, -8 singular
 *    matrix) go to status[p] and never affect other cells.
 *  - Contexts are not thread-safe; use one per host thread.
 */
#ifndef RADM2_H
#define RADM2_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RADM2_NVAR 59   /* variable species (mgpu.cu:18) */
#define RADM2_NFIX 2    /* fixed species H2O, M (mgpu.cu:20, 185-189) */
#define RADM2_NREACT 156 /* reactions (mgpu.cu:21) */
#define RADM2_NJ 21     /* photolysis inputs used by the rate constants */
#define RADM2_LU_NNZ 659 /* LU nonzeros (mgpu.cu:25) */

/* call-level error codes (KPP's -1..-5 are returned as is) */
#define RADM2_OK 0
#define RADM2_E_ARG (-101)         /* null pointer, bad size or stride */
#define RADM2_E_HIP (-102)         /* HIP runtime error */
#define RADM2_E_UNSUPPORTED (-103) /* option outside what the solver implements */

/* per-cell status codes (KPP IERR, mgpu.cu:976-1005) */
#define RADM2_CELL_OK 1
#define RADM2_CELL_TOO_MANY_STEPS (-6)
#define RADM2_CELL_STEP_TOO_SMALL (-7)
#define RADM2_CELL_SINGULAR (-8)

/* schedules */
#define RADM2_SCHED_GRID 0       /* one-shot grid: one lane per cell */
#define RADM2_SCHED_PERSISTENT 1 /* resident waves, work stealing over cells */

/*
 * Integration options, mirroring RPAR[0..6] and IPAR[2..3] of the
 * reference's Rosenbrock() (mgpu.cu:3698-3791): a zero field selects KPP's
 * default (Hmin 0, Hmax dt, Hstart max(Hmin,1e-6), FacMin 0.2, FacMax 6,
 * FacRej 0.1, FacSafe 0.9, 100000 steps); negative values are errors.
 * method: KPP's IPAR[3] number, 0 or 2 = Ros3 (the only method the
 * reference initialises), 1 Ros2, 3 Ros4, 4 Rodas3, 5 Rodas4.
 * sing_retries: step halvings after a singular matrix before -8 (KPP's
 * ros_PrepareMatrix allows 5; 0 = the reference's active code, mgpu.cu:
 * 3147-3163, which fails at once); negative values are RADM2_E_ARG.
 * Passing NULL instead selects INTEGRATE()'s call (mgpu.cu:3866-3873):
 * all defaults except Hstart = 0.01*dt.
 */
typedef struct radm2_opts {
  double hmin, hmax, hstart, facmin, facmax, facrej, facsafe;
  int32_t max_steps, method;
  int32_t sing_retries;
} radm2_opts;

typedef struct radm2_ctx radm2_ctx;
typedef struct radm2_plan radm2_plan;

/* One batch member for radm2_integrate_batched / radm2_plan_create. */
typedef struct radm2_domain {
  int64_t ncells, ld;
  double *conc;           /* [59][ld] concentrations (molec/cm3), in -> out */
  const double *temp;     /* [ld] temperature (K) */
  const double *cair;     /* [ld] air number density M (molec/cm3) */
  const double *h2o;      /* [ld] water (molec/cm3) */
  const double *jrate;    /* [21][ld] photolysis rates (s^-1) */
  const double *rc_n2o5;  /* [ld] N2O5 hydrolysis rate, NULL = reference value at RH 0 */
  const double *h0;       /* [ld] per-cell Hstart (entries <= 0 and NULL = the options') */
  int32_t *nstep, *nacc, *nrej, *status; /* [ld] per-cell outputs, each nullable */
  double *hexit;          /* [ld] step size at exit (nullable) */
  double *hfirst;         /* [ld] controller proposal after the first accepted step (nullable) */
  double dt, rtol, atol;
  const radm2_opts *opts; /* NULL = INTEGRATE() defaults */
  /* [21][ld] photolysis rates at the end of the interval (s^-1), or NULL.
   * Non-NULL makes the integration non-autonomous (KPP's IPAR[0] = 0): the
   * photolysis rate constants follow J(t) = jrate + (t/dt)(jrate_end -
   * jrate), dF/dT enters every stage (mgpu.cu:1566-1593, 3619-3622). */
  const double *jrate_end;
} radm2_domain;

/* Create a context on HIP device `device`.  stream: a hipStream_t to run on,
 * or NULL for a context-owned stream.  Replaces the reference's device setup
 * (initializeWRFChemData, mgpu.cu:419-511) for the solver's own needs. */
int radm2_create(radm2_ctx **ctx, int device, void *stream);
void radm2_destroy(radm2_ctx *ctx);

/* Schedule for later calls: RADM2_SCHED_GRID or RADM2_SCHED_PERSISTENT,
 * block size (64 or 256; 0 = tuned default), persistent refill threshold in
 * idle lanes per wave (1..64; 0 = tuned default).  No reference counterpart
 * (its launch shapes are hard-coded, e.g. mgpu.cu:1402-1404, 2797). */
int radm2_set_schedule(radm2_ctx *ctx, int schedule, int block, int refill_min);

/* Kernel variant for later calls (the autotuner's placement/occupancy axes):
 * placement of the lane-private working arrays -- RADM2_PLACE_PRIVATE
 * (compiler-managed registers, spills to scratch), RADM2_PLACE_WORKSPACE
 * (an explicit wave-interleaved HBM workspace whose loads are issued in
 * batches ahead of their use), RADM2_PLACE_LDS160 / RADM2_PLACE_LDS312 (the
 * workspace with its 160 / 312 most accessed doubles per lane held in LDS),
 * RADM2_PLACE_LDS160_REG112 (160 in LDS and the next 112 in registers; the
 * default; RADM2_PLACE_LDS160_REG128 is its former name, from when the
 * register tier held 128) -- the LDS placements run 64-lane blocks only, other block sizes
 * give RADM2_E_ARG at launch;
 * -1 = tuned default -- and the register budget as the minimum resident
 * waves per SIMD (1, 2, 4 or 8; 0 = tuned default).  Returns
 * RADM2_E_UNSUPPORTED for a combination that is not compiled in.  No
 * reference counterpart. */
#define RADM2_PLACE_PRIVATE 0
#define RADM2_PLACE_WORKSPACE 1
#define RADM2_PLACE_LDS160 2
#define RADM2_PLACE_LDS312 3
#define RADM2_PLACE_LDS160_REG112 4
#define RADM2_PLACE_LDS160_REG128 RADM2_PLACE_LDS160_REG112
/* A cell spread over 2 (SPLIT2) or 4 (SPLIT4) lanes that all run its
 * computation: the workspace plus, per cell, 160 / 320 doubles in LDS and
 * 256 / 512 in registers split over the cell's lanes (persistent schedule,
 * 64-lane blocks; the grid schedule gives RADM2_E_UNSUPPORTED). */
#define RADM2_PLACE_SPLIT2 5
#define RADM2_PLACE_SPLIT4 6
int radm2_set_kernel(radm2_ctx *ctx, int placement, int waves_per_simd);

/*
 * Integrate ncells cells over [0, dt] with rtol/atol for every species.
 * Replaces Rosenbrock() + RosenbrockIntegrator() (mgpu.cu:3664-3848,
 * 3412-3662) as called by INTEGRATE() (mgpu.cu:3850-3889), with the rate
 * constants of dev_Update_RCONST (mgpu.cu:752-919) evaluated inside the
 * kernel from temp/cair/jrate/rc_n2o5, and the cell state passed explicitly
 * instead of through globals.  Outputs: conc (end state), per-cell step,
 * accepted and rejected counts, status and exit step size.
 */
int radm2_integrate(radm2_ctx *ctx, int64_t ncells, int64_t ld, double *conc,
                    const double *temp, const double *cair, const double *h2o,
                    const double *jrate, const double *rc_n2o5, double dt, double rtol,
                    double atol, const radm2_opts *opts, const double *h0, int32_t *nstep,
                    int32_t *nacc, int32_t *nrej, int32_t *status, double *hexit);

/*
 * radm2_integrate without any device allocation: the persistent kernel's
 * workspace, its work counter and the domain descriptor live in a
 * file-scope device array of the library (128 MiB per device, held from the
 * first radm2_create on that device until the process ends), and the call
 * issues kernels only (no hipMalloc/hipFree, no copy from host memory, no
 * host synchronisation).  No counterpart in the reference: it is for
 * callers that cannot allocate -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.
 * Same arguments, semantics and results as radm2_integrate, always on the
 * persistent schedule (the context's kernel variant; the context's block
 * size when its schedule is persistent, else 64).  The array bounds the
 * resident waves, not the cell count: at most 236 waves of the Ros3
 * kernels (194 of the family kernel) run at once, about half of what
 * radm2_integrate uses on MI355X, so large batches take longer.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams, other contexts, other host threads) race and corrupt each other.
 * Returns RADM2_E_UNSUPPORTED if one block of the configured variant does
 * not fit the array.
 */
int radm2_integrate_noalloc(radm2_ctx *ctx, int64_t ncells, int64_t ld, double *conc,
                            const double *temp, const double *cair, const double *h2o,
                            const double *jrate, const double *rc_n2o5, double dt, double rtol,
                            double atol, const radm2_opts *opts, const double *h0,
                            int32_t *nstep, int32_t *nacc, int32_t *nrej, int32_t *status,
                            double *hexit);

/* Several independent domains (own arrays, dt, tolerances) in one launch.
 * No reference counterpart beyond its single hard-coded 39x19x39 domain
 * (mgpu.cu:420, 651-653); each domain follows radm2_integrate's semantics. */
int radm2_integrate_batched(radm2_ctx *ctx, int ndomains, const radm2_domain *doms);

/* Pre-validated batch captured as a HIP graph (counter reset + persistent
 * launch) for repeated calls on the same buffers. */
int radm2_plan_create(radm2_ctx *ctx, int ndomains, const radm2_domain *doms,
                      radm2_plan **plan);
int radm2_plan_launch(radm2_plan *plan);
void radm2_plan_destroy(radm2_plan *plan);

/* Wait for all work on the context's stream. */
int radm2_sync(radm2_ctx *ctx);

/* Device time in ms of the last integrate/launch call (HIP events on the
 * context's stream; valid after radm2_sync). Replaces the reference's
 * cudaStartTimer/cudaStopTimer (mgpu.cu:298-317). */
int radm2_last_elapsed_ms(radm2_ctx *ctx, float *ms);

/* Optional global counters summed over all cells of later calls: attempts,
 * accepted, rejected, persistent refill rounds.  The reference keeps global
 * Nstp/Nacc/Nrej (mgpu.cu:210, 3835-3842). */
int radm2_enable_stats(radm2_ctx *ctx, int on);
int radm2_read_stats(radm2_ctx *ctx, uint64_t out[4]);

/* Measured attainable FP64 throughput of the context's device for the
 * roofline report: eight independent FMA chains per lane in registers, the
 * best of three timed launches after a warm-up, in TFLOP/s.  Synchronizes
 * the context's stream; allocates and frees a 4 MB buffer.  No reference
 * counterpart: a measurement helper. */
int radm2_measure_fp64_peak(radm2_ctx *ctx, double *fp64_tflops);

/* KPP option check without running: returns 0, the IERR code -1..-5 of
 * Rosenbrock() (mgpu.cu:3715-3811), or RADM2_E_ARG for a negative
 * sing_retries. */
int radm2_check_opts(double dt, double rtol, double atol, const radm2_opts *opts);

/*
 * KPP front-end with the reference's own calling convention: Rosenbrock(
 * Tstart, Tend, AbsTol, RelTol, RPAR, IPAR) (mgpu.cu:3664-3848) over the
 * cells of `state` (its dt/rtol/atol/opts fields are ignored).  IPAR[0]
 * nonzero: autonomous (the reference's only use, mgpu.cu:3871); 0:
 * non-autonomous over state->jrate_end (NULL: J constant, so dF/dT = 0 and
 * the result is the autonomous one; Nfun counts the time derivative's
 * evaluation either way).  IPAR[1] selects scalar
 * (1) or vector (0) tolerances (vectors must hold one value for all 59
 * species), IPAR[2..3] and RPAR[0..6] as radm2_opts.  On return IPAR[10..17]
 * = Nfun, Njac, Nstp, Nacc, Nrej, Ndec, Nsol, Nsng summed over cells, and
 * RPAR[10..11] = Texit, Hexit (Tend and the largest exit step when every
 * cell succeeded).  Returns 1 (KPP success), -1..-8 or a RADM2_E_* code.
 * Synchronous.
 */
int radm2_rosenbrock(radm2_ctx *ctx, const radm2_domain *state, double Tstart, double Tend,
                     const double *AbsTol, const double *RelTol, double RPAR[20],
                     int32_t IPAR[20]);

/*
 * WRF-Chem fields -> the solver's inputs over a box of grid cells: the
 * reference's dev_convert (mgpu.cu:513-646, unit constants mgpu.cu:30-44).
 * Input arrays are WRF-ordered [field][ny][nz][nx] device arrays:
 *   phy   [3]:    pressure (Pa), temperature (K), air density (kg/m3)
 *   moist [>= 2]: field 1 = water vapour mixing ratio qv (kg/kg)
 *   ph    [>= 21]: photolysis rates (1/min), the reference's jv order
 *   chem  [>= 44], rad [>= 28]: mixing ratios (ppmv) in WRF's field order
 * Box cell (j, k, i), 0 <= j < by, 0 <= k < bz, 0 <= i < bx, reads input
 * (j + oy, k + oz, i + ox) (the reference: 39 x 19 x 39 cells at halo
 * offsets oy = ox = 4 of a 48 x 20 x 48 grid) and writes output cell
 * p = (j*bz + k)*bx + i of conc [59][ld], temp, cair, h2o, rc_n2o5 [ld]
 * and jrate [21][ld].  rc_fixed = 0 reproduces the reference's N2O5
 * hydrolysis rate as written (mgpu.cu:552, a misplaced parenthesis makes it
 * ~1e-135 s^-1); 1 uses 1/(3.6e4*exp(-(rh/0.28)^2.8) + 300).  Asynchronous
 * on `stream` (a hipStream_t; NULL = the default stream).
 */
typedef struct radm2_wrf_grid {
  int nx, ny, nz; /* input grid */
  int ox, oy, oz; /* box offset in the input grid */
  int bx, by, bz; /* box size */
} radm2_wrf_grid;
int radm2_wrf_convert(void *stream, const radm2_wrf_grid *grid, const double *phy,
                      const double *moist, const double *ph, const double *chem,
                      const double *rad, int rc_fixed, int64_t ld, double *conc, double *temp,
                      double *cair, double *h2o, double *jrate, double *rc_n2o5);

/* Text for a call-level or per-cell code (ros_ErrorMsg, mgpu.cu:976-1005). */
const char *radm2_strerror(int code);

#ifdef __cplusplus
}
#endif
#endif /* RADM2_H */
