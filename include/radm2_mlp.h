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

/* Mean squared error over the first `live` columns (surrogate heads). */
int radm2_mlp_mse_loss(void *stream, int M, int N, int live, const float *out, const float *tgt,
                       float *dout, float *loss);

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

/* h0[p] = clamp(10^out[p,0], 1e-12, hmax) if sigmoid(out[p,1]) > threshold,
 * else fallback (0 = the options' Hstart); used[p] (nullable) = 1 if taken. */
int radm2_mlp_h0_select(void *stream, int64_t n, int N, const float *out, double hmax,
                        float threshold, double fallback, double *h0, int32_t *used);

#ifdef __cplusplus
}
#endif
#endif /* RADM2_MLP_H */
