/*
 * radm2_mlp.h -- C ABI of the learned parts: BF16-MFMA / FP32-accumulate
 * dense-layer kernels and the pieces of their training, plus cell feature
 * extraction and the learned initial-step selection.
 *
 * No reference counterpart: the reference (mgpu.cu) has no learned parts.
 * These entry points serve the learned step-size proposal (h0, fed to
 * radm2_integrate's h0 argument) and the end-state surrogate, both judged by
 * the FP64 Rosenbrock path.
 *
 * Conventions: device pointers, row-major matrices, BF16 as uint16_t bit
 * patterns (round to nearest even), every GEMM dimension a multiple of 16.
 * `stream` is a hipStream_t (NULL = the default stream); all calls are
 * asynchronous.  Returns 0, or RADM2_E_ARG / RADM2_E_HIP (include/radm2.h).
 */
#ifndef RADM2_MLP_H
#define RADM2_MLP_H
#include <stdint.h>

#include "radm2.h"

#ifdef __cplusplus
extern "C" {
#endif

#define RADM2_ACT_NONE 0
#define RADM2_ACT_SILU 1
/* features per cell: log10 of 59 concentrations, T, log M, log H2O, log of
 * 21 photolysis rates */
#define RADM2_MLP_NFEAT 83

/* Z = X W + b (FP32, nullable), Y = act(Z) as BF16 (Yb) and/or FP32 (Yf).
 * X [M x K], W [K x N] BF16; b [N] FP32 (nullable). */
int radm2_mlp_dense_fwd(void *stream, int M, int K, int N, const uint16_t *X, const uint16_t *W,
                        const float *b, int act, float *Z, uint16_t *Yb, float *Yf);

/* dZ = dY * act'(Z): BF16 (dZb, the backward GEMMs' operand) and FP32 (dZf). */
int radm2_mlp_act_bwd(void *stream, int64_t n, const float *dY, const float *Z, int act,
                      uint16_t *dZb, float *dZf);

/* dX = dZ W^T (overwritten), dW += X^T dZ, db += column sums of dZf; each
 * output nullable. */
int radm2_mlp_dense_bwd(void *stream, int M, int K, int N, const uint16_t *X, const uint16_t *W,
                        const uint16_t *dZb, const float *dZf, float *dX, float *dW, float *db);

/* h0 head loss over rows of a [M x N] output: squared error of out[:,0]
 * against tgt[:,0] weighted by mask[:,0], binary cross-entropy of logit
 * out[:,1] against tgt[:,1] weighted by mask[:,1]; dout = d(mean)/d(out);
 * *loss += sum. */
int radm2_mlp_h0_loss(void *stream, int M, int N, const float *out, const float *tgt,
                      const float *mask, float *dout, float *loss);

/* Squared error of `heads` prediction blocks of one [M x N] output against
 * tgt [M x live]: head h is columns [h*width, h*width + live), weighted per
 * row by mask[m*heads + h] (mask NULL = 1); dout = d(sum/(M*live))/d(out),
 * 0 in padding columns; *loss += weighted sum.  live <= width,
 * heads*width <= N. */
int radm2_mlp_mse_loss(void *stream, int M, int N, int heads, int width, int live,
                       const float *out, const float *tgt, const float *mask, float *dout,
                       float *loss);

/* Adam step (1-based `step`) on FP32 weights w, refreshing the BF16 copy wb
 * (nullable); wd = decoupled-from-nothing L2 weight decay added to g. */
int radm2_mlp_adam(void *stream, int64_t n, float *w, uint16_t *wb, const float *g, float *m1,
                   float *m2, float lr, float b1, float b2, float eps, int step, float wd);

int radm2_mlp_to_bf16(void *stream, int64_t n, const float *x, uint16_t *y);

/* Standardised cell features [n x F] (F >= 83, zero padded) from solver-
 * layout inputs (SoA, cell index fastest). */
int radm2_mlp_features(void *stream, int64_t n, int64_t ld, const double *conc,
                       const double *temp, const double *cair, const double *h2o,
                       const double *jrate, const float *mu, const float *sd, int F, uint16_t *X,
                       float *Xf);

/* h0[p] = clamp(10^out[p,0], hlo, hmax) if sigmoid(out[p,1]) > threshold,
 * else fallback (0 = the options' Hstart); used[p] (nullable) = 1 if taken.
 * 0 < hlo <= hmax, else RADM2_E_ARG. */
int radm2_mlp_h0_select(void *stream, int64_t n, int N, const float *out, double hlo,
                        double hmax, float threshold, double fallback, double *h0,
                        int32_t *used);

/* Surrogate targets tgt[p*59 + s] = (log10 max(c1,floor) - log10 max(c0,floor)
 * - tmu[s]) / tsd[s] (tmu, tsd both NULL = unstandardised); c0, c1 [59][ld]. */
int radm2_mlp_dlog_targets(void *stream, int64_t n, int64_t ld, const double *c0, const double *c1,
                           double floor_c, const float *tmu, const float *tsd, float *tgt);

/* Surrogate judge over a two-head output [n x N] (head B at column `width`):
 * flag[p] = 1 (cell left untouched for the FP64 re-solve) when the heads'
 * largest disagreement in log10 space, over species predicted above cmin,
 * exceeds thr, or when the cell falls in the audit sample (fraction
 * `audit`, a hash of p and seed); otherwise conc[.][p] becomes the heads'
 * mean prediction and flag[p] = 0.  dis[p] (nullable) = the disagreement. */
int radm2_mlp_surrogate_apply(void *stream, int64_t n, int64_t ld, int N, int width,
                              const float *out, const float *tmu, const float *tsd, double floor_c,
                              double cmin, float thr, double audit, uint64_t seed, double *conc,
                              int32_t *flag, float *dis);

/* idx[0 .. *count) = indices of the cells with flag != 0 (unspecified order);
 * *count is reset on the stream first. */
int radm2_mlp_compact(void *stream, int64_t n, const int32_t *flag, int32_t *idx, int32_t *count);

/* Column gather (dst[r][i] = src[r][idx[i]]) and scatter (dst[r][idx[i]] =
 * src[r][i]) of `rows` SoA rows, i < m; elem = 4 (int32) or 8 (double). */
int radm2_mlp_gather(void *stream, int elem, int rows, int64_t lds, const void *src, int64_t m,
                     const int32_t *idx, int64_t ldd, void *dst);
int radm2_mlp_scatter(void *stream, int elem, int rows, int64_t lds, const void *src, int64_t m,
                      const int32_t *idx, int64_t ldd, void *dst);

#ifdef __cplusplus
}
#endif
#endif /* RADM2_MLP_H */
