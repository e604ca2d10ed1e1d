// Small-MLP kernels for the learned step-size proposal and end-state
// surrogate: BF16 operands on gfx950's MFMA (v_mfma_f32_16x16x32_bf16, wave64)
// with FP32 accumulation, plus the elementwise pieces of training
// (activation derivative, loss gradient, Adam) and the cell feature
// extraction.  No reference counterpart (the reference has no learned parts).
//
// Matrices are row-major.  The three products of a dense layer (forward,
// data gradient, weight gradient) run through one tiled kernel: a block of
// four waves computes a 64 x 64 output tile, staging 64 x 32 slices of both
// operands in LDS per step of the reduction (coalesced global loads along
// whichever index is contiguous, zero-filled outside the matrices), each
// wave a 32 x 32 quarter as 2 x 2 MFMA tiles.  Both operands sit in LDS
// with the reduction index contiguous, so a lane's fragment is one 16-byte
// read.
//
// MFMA 16x16x32 bf16 fragment layout (wave64), lane l:
//   A (16x32, rows x red):  A[l % 16][8*(l/16) + j], j = 0..7
//   B (32x16, red x cols):  B[8*(l/16) + j][l % 16], j = 0..7
//   D (16x16, rows x cols): D[4*(l/16) + j][l % 16], j = 0..3
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "radm2_mlp.h"

namespace radm2 {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
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

// C[M x N] = sum_r A(i, r) B(r, j) over r in [z*rchunk, min(R, (z+1)*rchunk)),
// with A(i, r) = A[i*ai + r*ar] and B(r, j) = B[r*br + j*bj].  Epilogues:
//   EPI_FWD     z = C + b, Z = z (FP32), Y = act(z) as BF16 (Yb) and/or FP32 (Yf)
//   EPI_STORE   C[i*ldc + j] = sum
//   EPI_ATOMIC  C[i*ldc + j] += sum (one FP32 atomic per element per chunk)
constexpr int GBM = 64, GBN = 64, GBK = 32, GPAD = 8;
enum { EPI_FWD = 0, EPI_STORE = 1, EPI_ATOMIC = 2 };

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int R, int rchunk,
                                              const uint16_t* __restrict__ A, int64_t ai, int64_t ar,
                                              const uint16_t* __restrict__ B, int64_t br, int64_t bj,
                                              float* __restrict__ C, int64_t ldc,
                                              const float* __restrict__ bias, int act,
                                              float* __restrict__ Z, uint16_t* __restrict__ Yb,
                                              float* __restrict__ Yf) {
  // reduction index contiguous in both tiles: A as [i][r], B as [j][r]
  __shared__ __attribute__((aligned(16))) uint16_t As[GBM][GBK + GPAD];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[GBN][GBK + GPAD];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int i0 = blockIdx.x * GBM, j0 = blockIdx.y * GBN;
  const int rb = blockIdx.z * rchunk, re = min(R, rb + rchunk);
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int u = 0; u < 2; u++) acc[s][u] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int r0 = rb; r0 < re; r0 += GBK) {
#pragma unroll
    for (int e = t; e < GBM * GBK; e += 256) {
      const int ii = ar == 1 ? e / GBK : e % GBM, rr = ar == 1 ? e % GBK : e / GBM;
      const int gi = i0 + ii, gr = r0 + rr;
      As[ii][rr] = (gi < M && gr < re) ? A[(int64_t)gi * ai + (int64_t)gr * ar] : (uint16_t)0;
    }
#pragma unroll
    for (int e = t; e < GBN * GBK; e += 256) {
      const int jj = bj == 1 ? e % GBN : e / GBK, rr = bj == 1 ? e / GBN : e % GBK;
      const int gj = j0 + jj, gr = r0 + rr;
      Bs[jj][rr] = (gj < N && gr < re) ? B[(int64_t)gr * br + (int64_t)gj * bj] : (uint16_t)0;
    }
    __syncthreads();
    bf16x8 fa[2], fb[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
      fa[s] = *reinterpret_cast<const bf16x8*>(&As[wm + 16 * s + (lane & 15)][8 * (lane >> 4)]);
      fb[s] = *reinterpret_cast<const bf16x8*>(&Bs[wn + 16 * s + (lane & 15)][8 * (lane >> 4)]);
    }
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
      for (int u = 0; u < 2; u++)
        acc[s][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[s], fb[u], acc[s][u], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int s = 0; s < 2; s++)
#pragma unroll
    for (int u = 0; u < 2; u++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int i = i0 + wm + 16 * s + 4 * (lane >> 4) + q, j = j0 + wn + 16 * u + (lane & 15);
        if (i >= M || j >= N) continue;
        const float v = acc[s][u][q];
        const int64_t o = (int64_t)i * ldc + j;
        if constexpr (EPI == EPI_FWD) {
          const float z = v + (bias ? bias[j] : 0.0f);
          const float y = act == RADM2_ACT_SILU ? silu(z) : z;
          if (Z) Z[o] = z;
          if (Yb) Yb[o] = f2bf(y);
          if (Yf) Yf[o] = y;
        } else if constexpr (EPI == EPI_STORE) {
          C[o] = v;
        } else {
          atomicAdd(&C[o], v);
        }
      }
}

static dim3 gemm_grid(int M, int N, int z) {
  return dim3((unsigned)((M + GBM - 1) / GBM), (unsigned)((N + GBN - 1) / GBN), (unsigned)z);
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
  // Y[M x N] = act(X[M x K] W[K x N] + b)
  hipLaunchKernelGGL(k_gemm<EPI_FWD>, gemm_grid(M, N, 1), dim3(256), 0, (hipStream_t)stream, M, N, K,
                     K, X, (int64_t)K, (int64_t)1, W, (int64_t)N, (int64_t)1, nullptr, (int64_t)N, b,
                     act, Z, Yb, Yf);
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
  // dX[M x K] = dZ[M x N] W^T (W is K x N)
  if (dX)
    hipLaunchKernelGGL(k_gemm<EPI_STORE>, gemm_grid(M, K, 1), dim3(256), 0, st, M, K, N, N, dZb,
                       (int64_t)N, (int64_t)1, W, (int64_t)1, (int64_t)N, dX, (int64_t)K, nullptr, 0,
                       nullptr, nullptr, nullptr);
  // dW[K x N] += X^T dZ, the batch rows in chunks of 1024 (one atomic add
  // per element per chunk)
  if (dW) {
    const int chunk = 1024;
    hipLaunchKernelGGL(k_gemm<EPI_ATOMIC>, gemm_grid(K, N, (M + chunk - 1) / chunk), dim3(256), 0, st,
                       K, N, M, chunk, X, (int64_t)1, (int64_t)K, dZb, (int64_t)N, (int64_t)1, dW,
                       (int64_t)N, nullptr, 0, nullptr, nullptr, nullptr);
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
