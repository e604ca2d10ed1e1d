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

// host-only table (a function-local static keeps it out of device code)
const KernelVariant* kernel_variants() {
#define RADM2_VARIANT_ROW(P, W, R)                                                        \
  {P, W, R, launch_grid_p##P##_w##W##_r##R, launch_persistent_p##P##_w##W##_r##R,          \
   persistent_blocks_per_cu_p##P##_w##W##_r##R}
  static const KernelVariant v[kNumVariants] = {
      RADM2_VARIANT_ROW(0, 1, 0), RADM2_VARIANT_ROW(0, 2, 0), RADM2_VARIANT_ROW(0, 4, 0),
      RADM2_VARIANT_ROW(0, 8, 0),
      RADM2_VARIANT_ROW(1, 1, 0), RADM2_VARIANT_ROW(1, 2, 0), RADM2_VARIANT_ROW(1, 4, 0),
      RADM2_VARIANT_ROW(2, 1, 0), RADM2_VARIANT_ROW(3, 1, 0), RADM2_VARIANT_ROW(4, 1, 0),
  };
#undef RADM2_VARIANT_ROW
  return v;
}

}  // namespace radm2
