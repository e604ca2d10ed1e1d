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

// Squared error of `heads` prediction blocks against one target row: head
// h occupies columns [h*width, h*width + live) of each N-wide output row and
// is weighted per row by mask[m*heads + h] (mask NULL = weight 1).  Writes
// dOut = d(sum / (M*live))/d(out) (zero in padding columns) and adds the
// weighted sum of squares to *loss.
__global__ void k_mse_loss(int M, int N, int heads, int width, int live,
                           const float* __restrict__ out, const float* __restrict__ tgt,
                           const float* __restrict__ mask, float* __restrict__ dout,
                           float* __restrict__ loss) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < (int64_t)M * N) {
    const int n = (int)(i % N);
    const int64_t m = i / N;
    const int h = n / width, c = n % width;
    if (h < heads && c < live) {
      const float w = mask ? mask[m * heads + h] : 1.0f;
      const float e = out[i] - tgt[m * live + c];
      dout[i] = 2.0f * e * w / ((float)M * live);
      l = e * e * w;
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
// [hlo, hmax] when sigmoid(out[1]) > threshold, else `fallback` (0 selects
// the options' Hstart in the solver).
__global__ void k_h0_select(int64_t n, int N, const float* __restrict__ out, double hlo,
                            double hmax, float threshold, double fallback,
                            double* __restrict__ h0, int32_t* __restrict__ used) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float conf = 1.0f / (1.0f + __expf(-out[p * N + 1]));
  const bool use = conf > threshold;
  double h = pow(10.0, (double)out[p * N]);
  h = fmin(fmax(h, hlo), hmax);
  h0[p] = use ? h : fallback;
  if (used) used[p] = use ? 1 : 0;
}

// Surrogate targets: per cell and species, the change of log10 max(c, floor)
// over the integration, standardised as (d - tmu[s]) / tsd[s] (tmu/tsd NULL =
// raw).  c0, c1: [59][ld] SoA; tgt: [n x 59] row-major.
__global__ void k_dlog_targets(int64_t n, int64_t ld, const double* __restrict__ c0,
                               const double* __restrict__ c1, double floor_c,
                               const float* __restrict__ tmu, const float* __restrict__ tsd,
                               float* __restrict__ tgt) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  for (int s = 0; s < 59; s++) {
    const double d = log10(fmax(c1[s * ld + p], floor_c)) - log10(fmax(c0[s * ld + p], floor_c));
    float v = (float)d;
    if (tmu) v = (v - tmu[s]) / tsd[s];
    tgt[p * 59 + s] = v;
  }
}

// 64-bit mix (splitmix64 finaliser) -> uniform [0, 1) for the audit draw.
__device__ __forceinline__ double audit_u01(uint64_t p, uint64_t seed) {
  uint64_t z = p + seed * 0x9E3779B97F4A7C15ull + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

// Surrogate judge, per cell: head predictions dA, dB of the log10 change of
// each species (de-standardised); the heads' disagreement is the largest
// |dA - dB| over species whose predicted end value exceeds cmin.  A cell is
// flagged for the FP64 re-solve when the disagreement exceeds thr or when
// it falls in the audit sample (fraction audit, drawn from a hash of the
// cell index and seed); otherwise its concentrations are replaced by the
// heads' mean prediction.  Flagged cells keep their input state.
__global__ void k_surrogate_apply(int64_t n, int64_t ld, int N, int width,
                                  const float* __restrict__ out, const float* __restrict__ tmu,
                                  const float* __restrict__ tsd, double floor_c, double cmin,
                                  float thr, double audit, uint64_t seed,
                                  double* __restrict__ conc, int32_t* __restrict__ flag,
                                  float* __restrict__ dis) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float* o = out + p * N;
  float worst = 0.0f;
  for (int s = 0; s < 59; s++) {
    const float a = o[s] * tsd[s] + tmu[s], b = o[width + s] * tsd[s] + tmu[s];
    const double l0 = log10(fmax(conc[s * ld + p], floor_c));
    if (l0 + (double)fmaxf(a, b) > log10(cmin)) worst = fmaxf(worst, fabsf(a - b));
  }
  const bool f = !(worst <= thr) || audit_u01((uint64_t)p, seed) < audit;
  flag[p] = f ? 1 : 0;
  if (dis) dis[p] = worst;
  if (f) return;
  for (int s = 0; s < 59; s++) {
    const float a = o[s] * tsd[s] + tmu[s], b = o[width + s] * tsd[s] + tmu[s];
    const double l0 = log10(fmax(conc[s * ld + p], floor_c));
    conc[s * ld + p] = pow(10.0, l0 + 0.5 * ((double)a + (double)b));
  }
}

__global__ void k_zero_i32(int32_t* __restrict__ x) { *x = 0; }

// idx[0 .. *count) = the cells with flag != 0, in unspecified order
// (wave-aggregated: one atomic per wave).
__global__ void k_compact(int64_t n, const int32_t* __restrict__ flag, int32_t* __restrict__ idx,
                          int32_t* __restrict__ count) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool f = p < n && flag[p] != 0;
  const uint64_t mask = __ballot(f);
  const int lane = threadIdx.x & 63;
  int base = 0;
  if (lane == 0 && mask) base = atomicAdd(count, (int32_t)__popcll(mask));
  base = __shfl(base, 0);
  if (f) idx[base + __popcll(mask & ((1ull << lane) - 1ull))] = (int32_t)p;
}

// Column gather/scatter over SoA arrays of `rows` rows: gather dst[r][i] =
// src[r][idx[i]], scatter dst[r][idx[i]] = src[r][i], for i < m.
template <typename T>
__global__ void k_gather(int rows, int64_t lds, const T* __restrict__ src, int64_t m,
                         const int32_t* __restrict__ idx, int64_t ldd, T* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int64_t c = idx[i];
  for (int r = 0; r < rows; r++) dst[r * ldd + i] = src[r * lds + c];
}
template <typename T>
__global__ void k_scatter(int rows, int64_t lds, const T* __restrict__ src, int64_t m,
                          const int32_t* __restrict__ idx, int64_t ldd, T* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int64_t c = idx[i];
  for (int r = 0; r < rows; r++) dst[r * ldd + c] = src[r * lds + i];
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

int radm2_mlp_mse_loss(void* stream, int M, int N, int heads, int width, int live,
                       const float* out, const float* tgt, const float* mask, float* dout,
                       float* loss) {
  if (M <= 0 || N <= 0 || heads <= 0 || width <= 0 || live <= 0 || live > width ||
      (int64_t)heads * width > N || !out || !tgt || !dout || !loss)
    return RADM2_E_ARG;
  hipLaunchKernelGGL(k_mse_loss, g1((int64_t)M * N, 256), dim3(256), 0, (hipStream_t)stream, M, N,
                     heads, width, live, out, tgt, mask, dout, loss);
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

int radm2_mlp_h0_select(void* stream, int64_t n, int N, const float* out, double hlo,
                        double hmax, float threshold, double fallback, double* h0,
                        int32_t* used) {
  if (n <= 0 || N < 2 || !out || !h0 || !(hlo > 0.0) || !(hmax >= hlo)) return RADM2_E_ARG;
  hipLaunchKernelGGL(k_h0_select, g1(n, 256), dim3(256), 0, (hipStream_t)stream, n, N, out, hlo,
                     hmax, threshold, fallback, h0, used);
  return launch_ok();
}

int radm2_mlp_dlog_targets(void* stream, int64_t n, int64_t ld, const double* c0, const double* c1,
                           double floor_c, const float* tmu, const float* tsd, float* tgt) {
  if (n <= 0 || ld < n || !c0 || !c1 || !tgt || !(floor_c > 0.0) || (!tmu != !tsd))
    return RADM2_E_ARG;
  hipLaunchKernelGGL(k_dlog_targets, g1(n, 128), dim3(128), 0, (hipStream_t)stream, n, ld, c0, c1,
                     floor_c, tmu, tsd, tgt);
  return launch_ok();
}

int radm2_mlp_surrogate_apply(void* stream, int64_t n, int64_t ld, int N, int width,
                              const float* out, const float* tmu, const float* tsd, double floor_c,
                              double cmin, float thr, double audit, uint64_t seed, double* conc,
                              int32_t* flag, float* dis) {
  if (n <= 0 || ld < n || width < 59 || 2 * width > N || !out || !tmu || !tsd || !conc || !flag ||
      !(floor_c > 0.0) || !(cmin > 0.0) || !(thr >= 0.0f) || !(audit >= 0.0 && audit <= 1.0))
    return RADM2_E_ARG;
  hipLaunchKernelGGL(k_surrogate_apply, g1(n, 128), dim3(128), 0, (hipStream_t)stream, n, ld, N,
                     width, out, tmu, tsd, floor_c, cmin, thr, audit, seed, conc, flag, dis);
  return launch_ok();
}

int radm2_mlp_compact(void* stream, int64_t n, const int32_t* flag, int32_t* idx, int32_t* count) {
  if (n <= 0 || n > INT32_MAX || !flag || !idx || !count) return RADM2_E_ARG;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(1), 0, st, count);
  hipLaunchKernelGGL(k_compact, g1(n, 256), dim3(256), 0, st, n, flag, idx, count);
  return launch_ok();
}

int radm2_mlp_gather(void* stream, int elem, int rows, int64_t lds, const void* src, int64_t m,
                     const int32_t* idx, int64_t ldd, void* dst) {
  if (rows <= 0 || m < 0 || ldd < m || !src || !dst || (m > 0 && !idx) || (elem != 4 && elem != 8))
    return RADM2_E_ARG;
  if (m == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (elem == 8)
    hipLaunchKernelGGL(k_gather<double>, g1(m, 256), dim3(256), 0, st, rows, lds,
                       (const double*)src, m, idx, ldd, (double*)dst);
  else
    hipLaunchKernelGGL(k_gather<int32_t>, g1(m, 256), dim3(256), 0, st, rows, lds,
                       (const int32_t*)src, m, idx, ldd, (int32_t*)dst);
  return launch_ok();
}

int radm2_mlp_scatter(void* stream, int elem, int rows, int64_t lds, const void* src, int64_t m,
                      const int32_t* idx, int64_t ldd, void* dst) {
  if (rows <= 0 || m < 0 || lds < m || !src || !dst || (m > 0 && !idx) || (elem != 4 && elem != 8))
    return RADM2_E_ARG;
  if (m == 0) return 0;
  hipStream_t st = (hipStream_t)stream;
  if (elem == 8)
    hipLaunchKernelGGL(k_scatter<double>, g1(m, 256), dim3(256), 0, st, rows, lds,
                       (const double*)src, m, idx, ldd, (double*)dst);
  else
    hipLaunchKernelGGL(k_scatter<int32_t>, g1(m, 256), dim3(256), 0, st, rows, lds,
                       (const int32_t*)src, m, idx, ldd, (int32_t*)dst);
  return launch_ok();
}

}  // extern "C"
