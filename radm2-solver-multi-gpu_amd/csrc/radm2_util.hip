// Small helper kernels.
#include "radm2_kernels.h"

namespace radm2 {

__global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_u64, dim3((n + 63) / 64), dim3(64), 0, st, p, n, v);
  return hipGetLastError();
}

__global__ void k_slot_init(unsigned long long* bits, int nwords, int64_t nslots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nwords) return;
  const int64_t lo = (int64_t)i * 64;
  unsigned long long v = 0ull;
  if (nslots <= lo) v = ~0ull;
  else if (nslots < lo + 64) v = ~0ull << (nslots - lo);
  bits[i] = v;
}

hipError_t launch_slot_init(unsigned long long* bits, int nwords, int64_t nslots, hipStream_t st) {
  hipLaunchKernelGGL(k_slot_init, dim3((nwords + 63) / 64), dim3(64), 0, st, bits, nwords, nslots);
  return hipGetLastError();
}

// Attainable FP64 peak for the roofline report (measurement only): eight
// independent FMA chains per lane held in registers; the lane's sum is
// stored so that no chain is dead.  b < 1 keeps the chains bounded.
__global__ void __launch_bounds__(256) k_peak_fma(double* out, int iters, double b, double c) {
  double a0 = 1e-3 * threadIdx.x, a1 = a0 + 1.0, a2 = a0 + 2.0, a3 = a0 + 3.0;
  double a4 = a0 + 4.0, a5 = a0 + 5.0, a6 = a0 + 6.0, a7 = a0 + 7.0;
  for (int it = 0; it < iters; it++) {
    a0 = fma(a0, b, c);
    a1 = fma(a1, b, c);
    a2 = fma(a2, b, c);
    a3 = fma(a3, b, c);
    a4 = fma(a4, b, c);
    a5 = fma(a5, b, c);
    a6 = fma(a6, b, c);
    a7 = fma(a7, b, c);
  }
  out[blockIdx.x * blockDim.x + threadIdx.x] = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
}

hipError_t measure_fp64_peak(hipStream_t st, int cus, double* fp64_tflops) {
  constexpr int kThreads = 256, kIters = 200000;
  const int blocks = cus * 8;  // 8 waves per SIMD
  double* out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t err = hipSuccess;
  float best = 0.0f;
  if ((err = hipMalloc(&out, sizeof(double) * blocks * kThreads)) != hipSuccess) goto done;
  if ((err = hipEventCreate(&e0)) != hipSuccess) goto done;
  if ((err = hipEventCreate(&e1)) != hipSuccess) goto done;
  // one warm-up launch, then the best of three
  for (int rep = 0; rep < 4; rep++) {
    float ms = 0.0f;
    if ((err = hipEventRecord(e0, st)) != hipSuccess) goto done;
    hipLaunchKernelGGL(k_peak_fma, dim3(blocks), dim3(kThreads), 0, st, out, kIters, 0.999999,
                       1e-7);
    if ((err = hipGetLastError()) != hipSuccess) goto done;
    if ((err = hipEventRecord(e1, st)) != hipSuccess) goto done;
    if ((err = hipEventSynchronize(e1)) != hipSuccess) goto done;
    if ((err = hipEventElapsedTime(&ms, e0, e1)) != hipSuccess) goto done;
    if (rep > 0 && (best == 0.0f || ms < best)) best = ms;
  }
  *fp64_tflops = 2.0 * 8.0 * kIters * double(blocks) * kThreads / (best * 1e-3) / 1e12;
done:
  (void)hipStreamSynchronize(st);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (out) (void)hipFree(out);
  return err;
}

// host-only table (a function-local static keeps it out of device code)
const KernelVariant* kernel_variants() {
#define RADM2_VARIANT_ROW_NAMED(P, W, R, F, S)                                              \
  {P, W, R, F, F ? WS_DOUBLES_FAMILY : WS_DOUBLES_ROS3, hot_lanes(P), launch_grid##S,         \
   launch_persistent##S, persistent_blocks_per_cu##S, grid_blocks_per_cu##S}
#define RADM2_VARIANT_ROW(P, W, R) RADM2_VARIANT_ROW_NAMED(P, W, R, 0, _p##P##_w##W##_r##R)
  static const KernelVariant v[kNumVariants] = {
      RADM2_VARIANT_ROW(0, 1, 0), RADM2_VARIANT_ROW(0, 2, 0), RADM2_VARIANT_ROW(0, 4, 0),
      RADM2_VARIANT_ROW(0, 8, 0),
      RADM2_VARIANT_ROW(1, 1, 0), RADM2_VARIANT_ROW(1, 2, 0), RADM2_VARIANT_ROW(1, 4, 0),
      RADM2_VARIANT_ROW(2, 1, 0), RADM2_VARIANT_ROW(3, 1, 0), RADM2_VARIANT_ROW(4, 1, 0),
      RADM2_VARIANT_ROW(5, 1, 0), RADM2_VARIANT_ROW(6, 1, 0),
      RADM2_VARIANT_ROW_NAMED(4, 1, 0, 1, _p4_w1_r0_m1),
  };
#undef RADM2_VARIANT_ROW
#undef RADM2_VARIANT_ROW_NAMED
  return v;
}

}  // namespace radm2
