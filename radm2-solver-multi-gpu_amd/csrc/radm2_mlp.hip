// Small-MLP kernels for the learned step-size proposal and end-state
// surrogate: BF16 operands on MFMA (v_mfma_f32_16x16x16_bf16, wave64) with
// FP32 accumulation, plus the elementwise pieces of training (activation
// derivative, loss gradient, Adam) and the cell feature extraction.
// No reference counterpart (the reference has no learned parts).
//
// Matrices are row-major.  Every dimension the GEMMs see is a multiple of
// 16 (the host pads features to 96 and outputs to 16); each wave computes
// one 16x16 output tile, walking the reduction dimension 16 at a time.
//
// MFMA 16x16x16 bf16 fragment layout (wave64), lane l, item j = 0..3:
//   A (16x16, rows x red):  A[l % 16][4*(l/16) + j]
//   B (16x16, red x cols):  B[4*(l/16) + j][l % 16]
//   D (16x16, rows x cols): D[4*(l/16) + j][l % 16]
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "radm2_mlp.h"

namespace radm2 {
namespace mlp {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }
__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// Y[M x N] = act(X[M x K] W[K x N] + b); Z keeps the pre-activation (FP32)
// for the backward pass; Y in BF16 (next layer's operand) and/or FP32.
__global__ __launch_bounds__(64) void k_dense_fwd(int M, int K, int N, const uint16_t* __restrict__ X,
                                                  const uint16_t* __restrict__ W,
                                                  const float* __restrict__ b, int act,
                                                  float* __restrict__ Z, uint16_t* __restrict__ Yb,
                                                  float* __restrict__ Yf) {
  const int lane = threadIdx.x;
  const int m0 = blockIdx.x * 16, n0 = blockIdx.y * 16;
  const int r = lane & 15, q = lane >> 4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    bf16x4 a, bb;
    for (int j = 0; j < 4; j++) {
      a[j] = (short)X[(size_t)(m0 + r) * K + k0 + 4 * q + j];
      bb[j] = (short)W[(size_t)(k0 + 4 * q + j) * N + n0 + r];
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
  }
  for (int j = 0; j < 4; j++) {
    const int m = m0 + 4 * q + j, n = n0 + r;
    const float z = acc[j] + (b ? b[n] : 0.0f);
    const float y = act == RADM2_ACT_SILU ? silu(z) : z;
    const size_t o = (size_t)m * N + n;
    if (Z) Z[o] = z;
    if (Yb) Yb[o] = f2bf(y);
    if (Yf) Yf[o] = y;
  }
}

// dZ = dY * act'(Z) as BF16 (the backward GEMMs' operand); FP32 copy kept
// for the bias gradient.
__global__ void k_act_bwd(int64_t n, const float* __restrict__ dY, const float* __restrict__ Z,
                          int act, uint16_t* __restrict__ dZb, float* __restrict__ dZf) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = act == RADM2_ACT_SILU ? dY[i] * silu_grad(Z[i]) : dY[i];
  dZb[i] = f2bf(g);
  dZf[i] = g;
}

// dX[M x K] = dZ[M x N] W^T  (W is K x N)
__global__ __launch_bounds__(64) void k_dense_bwd_data(int M, int K, int N,
                                                       const uint16_t* __restrict__ dZ,
                                                       const uint16_t* __restrict__ W,
                                                       float* __restrict__ dX) {
  const int lane = threadIdx.x;
  const int m0 = blockIdx.x * 16, k0 = blockIdx.y * 16;
  const int r = lane & 15, q = lane >> 4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int n1 = 0; n1 < N; n1 += 16) {
    bf16x4 a, bb;
    for (int j = 0; j < 4; j++) {
      a[j] = (short)dZ[(size_t)(m0 + r) * N + n1 + 4 * q + j];       // A[m][n]
      bb[j] = (short)W[(size_t)(k0 + r) * N + n1 + 4 * q + j];       // B[n][k] = W[k][n]
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
  }
  for (int j = 0; j < 4; j++) dX[(size_t)(m0 + 4 * q + j) * K + k0 + r] = acc[j];
}

// dW[K x N] += X^T dZ over rows [mb*chunk, (mb+1)*chunk) of the batch, one
// FP32 atomic add per element per chunk.
__global__ __launch_bounds__(64) void k_dense_bwd_weight(int M, int K, int N, int chunk,
                                                         const uint16_t* __restrict__ X,
                                                         const uint16_t* __restrict__ dZ,
                                                         float* __restrict__ dW) {
  const int lane = threadIdx.x;
  const int k0 = blockIdx.x * 16, n0 = blockIdx.y * 16;
  const int mb = blockIdx.z * chunk, me = min(M, mb + chunk);
  const int r = lane & 15, q = lane >> 4;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int m1 = mb; m1 < me; m1 += 16) {
    bf16x4 a, bb;
    for (int j = 0; j < 4; j++) {
      a[j] = (short)X[(size_t)(m1 + 4 * q + j) * K + k0 + r];        // A[k][m] = X[m][k]
      bb[j] = (short)dZ[(size_t)(m1 + 4 * q + j) * N + n0 + r];      // B[m][n]
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
  }
  for (int j = 0; j < 4; j++) atomicAdd(&dW[(size_t)(k0 + 4 * q + j) * N + n0 + r], acc[j]);
}

// db[N] += sum over rows of dZ (FP32)
__global__ void k_bias_grad(int M, int N, const float* __restrict__ dZ, float* __restrict__ db) {
  const int n = blockIdx.x;
  float s = 0.0f;
  for (int m = threadIdx.x; m < M; m += blockDim.x) s += dZ[(size_t)m * N + n];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
  __shared__ float part[16];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0.0f;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) t += part[w];
    atomicAdd(&db[n], t);
  }
}

// Loss on the 2 live outputs of a 16-wide output row:
//   out[:,0] = predicted log10 h0, regressed on target[:,0] with weight
//   mask[:,0] (squared error), out[:,1] = confidence logit, binary cross
//   entropy against target[:,1] in {0,1} with weight mask[:,1].
// Writes dOut (mean over rows) and accumulates the loss sum.
__global__ void k_h0_loss(int M, int N, const float* __restrict__ out, const float* __restrict__ tgt,
                          const float* __restrict__ mask, float* __restrict__ dout,
                          float* __restrict__ loss) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (m < M) {
    const float* o = out + (size_t)m * N;
    float* d = dout + (size_t)m * N;
    for (int n = 0; n < N; n++) d[n] = 0.0f;
    const float e = o[0] - tgt[2 * m];
    d[0] = 2.0f * e * mask[2 * m] / M;
    l += e * e * mask[2 * m];
    const float p = 1.0f / (1.0f + __expf(-o[1]));
    d[1] = (p - tgt[2 * m + 1]) * mask[2 * m + 1] / M;
    const float t = tgt[2 * m + 1];
    l += -(t * __logf(fmaxf(p, 1e-7f)) + (1.0f - t) * __logf(fmaxf(1.0f - p, 1e-7f))) * mask[2 * m + 1];
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
}

// Mean squared error over the first `live` columns of each row (surrogate
// heads): dOut and the loss sum.
__global__ void k_mse_loss(int M, int N, int live, const float* __restrict__ out,
                           const float* __restrict__ tgt, float* __restrict__ dout,
                           float* __restrict__ loss) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < (int64_t)M * N) {
    const int n = (int)(i % N);
    const int64_t m = i / N;
    if (n < live) {
      const float e = out[i] - tgt[m * live + n];
      dout[i] = 2.0f * e / ((float)M * live);
      l = e * e;
    } else {
      dout[i] = 0.0f;
    }
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
}

// Adam on FP32 master weights; refreshes the BF16 copy the GEMMs read.
__global__ void k_adam(int64_t n, float* __restrict__ w, uint16_t* __restrict__ wb,
                       const float* __restrict__ g, float* __restrict__ m1, float* __restrict__ m2,
                       float lr, float b1, float b2, float eps, float bc1, float bc2, float wd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * w[i];
  m1[i] = b1 * m1[i] + (1.0f - b1) * gi;
  m2[i] = b2 * m2[i] + (1.0f - b2) * gi * gi;
  const float mh = m1[i] / bc1, vh = m2[i] / bc2;
  w[i] -= lr * mh / (sqrtf(vh) + eps);
  if (wb) wb[i] = f2bf(w[i]);
}

__global__ void k_to_bf16(int64_t n, const float* __restrict__ x, uint16_t* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

// Cell features (row m of X, BF16, width F >= 83, zero padded):
//   log10 of the 59 concentrations (floor 1e-30), T/300, log10 M - 19,
//   log10 H2O - 17, log10(J_i + 1e-12) + 6 for the 21 photolysis rates,
// each then standardised with the given mean/scale (length 83).
__global__ void k_features(int64_t n, int64_t ld, const double* __restrict__ conc,
                           const double* __restrict__ temp, const double* __restrict__ cair,
                           const double* __restrict__ h2o, const double* __restrict__ jrate,
                           const float* __restrict__ mu, const float* __restrict__ sd, int F,
                           uint16_t* __restrict__ X, float* __restrict__ Xf) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float f[RADM2_MLP_NFEAT];
  for (int s = 0; s < 59; s++) f[s] = (float)log10(fmax(conc[s * ld + p], 1e-30));
  f[59] = (float)(temp[p] / 300.0);
  f[60] = (float)(log10(cair[p]) - 19.0);
  f[61] = (float)(log10(fmax(h2o[p], 1.0)) - 17.0);
  for (int j = 0; j < 21; j++) f[62 + j] = (float)(log10(jrate[j * ld + p] + 1e-12) + 6.0);
  for (int i = 0; i < F; i++) {
    const float v = i < RADM2_MLP_NFEAT ? (f[i] - mu[i]) / sd[i] : 0.0f;
    if (X) X[p * F + i] = f2bf(v);
    if (Xf) Xf[p * F + i] = v;
  }
}

// Learned initial step from the network outputs: use 10^out[0] clamped to
// [1e-12, hmax] when sigmoid(out[1]) > threshold, else `fallback` (0 selects
// the options' Hstart in the solver).
__global__ void k_h0_select(int64_t n, int N, const float* __restrict__ out, double hmax,
                            float threshold, double fallback, double* __restrict__ h0,
                            int32_t* __restrict__ used) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float conf = 1.0f / (1.0f + __expf(-out[p * N + 1]));
  const bool use = conf > threshold;
  double h = pow(10.0, (double)out[p * N]);
  h = fmin(fmax(h, 1.0e-12), hmax);
  h0[p] = use ? h : fallback;
  if (used) used[p] = use ? 1 : 0;
}

}  // namespace mlp
}  // namespace radm2

using namespace radm2::mlp;

static int launch_ok() { return hipGetLastError() == hipSuccess ? 0 : RADM2_E_HIP; }
static dim3 g1(int64_t n, int b) { return dim3((unsigned)((n + b - 1) / b)); }

extern "C" {

int radm2_mlp_dense_fwd(void* stream, int M, int K, int N, const uint16_t* X, const uint16_t* W,
                        const float* b, int act, float* Z, uint16_t* Yb, float* Yf) {
  if (M <= 0 || K <= 0 || N <= 0 || (M | K | N) & 15 || !X || !W) return RADM2_E_ARG;
  hipLaunchKernelGGL(k_dense_fwd, dim3(M / 16, N / 16), dim3(64), 0, (hipStream_t)stream, M, K, N,
                     X, W, b, act, Z, Yb, Yf);
  return launch_ok();
}

int radm2_mlp_act_bwd(void* stream, int64_t n, const float* dY, const float* Z, int act,
                      uint16_t* dZb, float* dZf) {
  if (n <= 0 || !dY || !dZb || !dZf || (act == RADM2_ACT_SILU && !Z)) return RADM2_E_ARG;
  hipLaunchKernelGGL(k_act_bwd, g1(n, 256), dim3(256), 0, (hipStream_t)stream, n, dY, Z, act, dZb,
                     dZf);
  return launch_ok();
}

int radm2_mlp_dense_bwd(void* stream, int M, int K, int N, const uint16_t* X, const uint16_t* W,
                        const uint16_t* dZb, const float* dZf, float* dX, float* dW, float* db) {
  if (M <= 0 || K <= 0 || N <= 0 || (M | K | N) & 15 || !X || !W || !dZb) return RADM2_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  if (dX) hipLaunchKernelGGL(k_dense_bwd_data, dim3(M / 16, K / 16), dim3(64), 0, st, M, K, N, dZb, W, dX);
  if (dW) {
    const int chunk = 1024;
    hipLaunchKernelGGL(k_dense_bwd_weight, dim3(K / 16, N / 16, (M + chunk - 1) / chunk), dim3(64), 0,
                       st, M, K, N, chunk, X, dZb, dW);
  }
  if (db) {
    if (!dZf) return RADM2_E_ARG;
    hipLaunchKernelGGL(k_bias_grad, dim3(N), dim3(256), 0, st, M, N, dZf, db);
  }
  return launch_ok();
}

int radm2_mlp_h0_loss(void* stream, int M, int N, const float* out, const float* tgt,
                      const float* mask, float* dout, float* loss) {
  if (M <= 0 || N < 2 || !out || !tgt || !mask || !dout || !loss) return RADM2_E_ARG;
  hipLaunchKernelGGL(k_h0_loss, g1(M, 256), dim3(256), 0, (hipStream_t)stream, M, N, out, tgt, mask,
                     dout, loss);
  return launch_ok();
}

int radm2_mlp_mse_loss(void* stream, int M, int N, int live, const float* out, const float* tgt,
                       float* dout, float* loss) {
  if (M <= 0 || N <= 0 || live <= 0 || live > N || !out || !tgt || !dout || !loss) return RADM2_E_ARG;
  hipLaunchKernelGGL(k_mse_loss, g1((int64_t)M * N, 256), dim3(256), 0, (hipStream_t)stream, M, N,
                     live, out, tgt, dout, loss);
  return launch_ok();
}

int radm2_mlp_adam(void* stream, int64_t n, float* w, uint16_t* wb, const float* g, float* m1,
                   float* m2, float lr, float b1, float b2, float eps, int step, float wd) {
  if (n <= 0 || !w || !g || !m1 || !m2 || step <= 0) return RADM2_E_ARG;
  const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, g1(n, 256), dim3(256), 0, (hipStream_t)stream, n, w, wb, g, m1, m2, lr,
                     b1, b2, eps, bc1, bc2, wd);
  return launch_ok();
}

int radm2_mlp_to_bf16(void* stream, int64_t n, const float* x, uint16_t* y) {
  if (n <= 0 || !x || !y) return RADM2_E_ARG;
  hipLaunchKernelGGL(k_to_bf16, g1(n, 256), dim3(256), 0, (hipStream_t)stream, n, x, y);
  return launch_ok();
}

int radm2_mlp_features(void* stream, int64_t n, int64_t ld, const double* conc, const double* temp,
                       const double* cair, const double* h2o, const double* jrate, const float* mu,
                       const float* sd, int F, uint16_t* X, float* Xf) {
  if (n <= 0 || ld < n || F < RADM2_MLP_NFEAT || !conc || !temp || !cair || !h2o || !jrate || !mu ||
      !sd || (!X && !Xf))
    return RADM2_E_ARG;
  hipLaunchKernelGGL(k_features, g1(n, 128), dim3(128), 0, (hipStream_t)stream, n, ld, conc, temp,
                     cair, h2o, jrate, mu, sd, F, X, Xf);
  return launch_ok();
}

int radm2_mlp_h0_select(void* stream, int64_t n, int N, const float* out, double hmax,
                        float threshold, double fallback, double* h0, int32_t* used) {
  if (n <= 0 || N < 2 || !out || !h0 || !(hmax > 0.0)) return RADM2_E_ARG;
  hipLaunchKernelGGL(k_h0_select, g1(n, 256), dim3(256), 0, (hipStream_t)stream, n, N, out, hmax,
                     threshold, fallback, h0, used);
  return launch_ok();
}

}  // extern "C"
