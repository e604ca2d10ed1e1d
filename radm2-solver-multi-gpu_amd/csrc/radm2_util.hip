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
#define RADM2_VARIANT_ROW_NAMED(P, W, R, F, S)                                              \
  {P, W, R, F, F ? WS_DOUBLES_FAMILY : WS_DOUBLES_ROS3, launch_grid##S, launch_persistent##S,   \
   persistent_blocks_per_cu##S}
#define RADM2_VARIANT_ROW(P, W, R) RADM2_VARIANT_ROW_NAMED(P, W, R, 0, _p##P##_w##W##_r##R)
  static const KernelVariant v[kNumVariants] = {
      RADM2_VARIANT_ROW(0, 1, 0), RADM2_VARIANT_ROW(0, 2, 0), RADM2_VARIANT_ROW(0, 4, 0),
      RADM2_VARIANT_ROW(0, 8, 0),
      RADM2_VARIANT_ROW(1, 1, 0), RADM2_VARIANT_ROW(1, 2, 0), RADM2_VARIANT_ROW(1, 4, 0),
      RADM2_VARIANT_ROW(2, 1, 0), RADM2_VARIANT_ROW(3, 1, 0), RADM2_VARIANT_ROW(4, 1, 0),
      RADM2_VARIANT_ROW_NAMED(4, 1, 0, 1, _p4_w1_r0_m1),
  };
#undef RADM2_VARIANT_ROW
#undef RADM2_VARIANT_ROW_NAMED
  return v;
}

}  // namespace radm2
