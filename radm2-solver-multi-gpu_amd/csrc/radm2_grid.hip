// One-shot grid kernel: one lane per cell for the whole integration.
// A wave runs until its slowest cell finishes (the reference's schedule,
// mgpu.cu:3504-3641, with the per-cell loop moved onto the lane).
#include "radm2_cell_io.h"

namespace radm2 {

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_ros3_grid(LaunchArgs a) {
  const int64_t g = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (g >= a.total) return;
  const int di = find_domain(a.doms, a.ndom, g);
  const DomainDesc& d = a.doms[di];
  const int64_t p = g - d.first;
  double kk[NREACT], y[NVAR], lu[LU_NNZ], f[NVAR], k1[NVAR], k2[NVAR], k3[NVAR];
  double h2o, m;
  CellState s;
  cell_load(d, p, kk, y, h2o, m, s);
  const Ros3Opts& o = d.opts;
  for (;;) {
    if (s.newstep) cell_step_top(s, o);
    if (s.status != ST_RUNNING) break;
    ros3_attempt(s, o, kk, y, h2o, m, lu, f, k1, k2, k3);
    if (s.status != ST_RUNNING) break;
  }
  cell_store(d, p, y, s, s.hfirst, a.stats);
}

hipError_t launch_grid(int block, const LaunchArgs& a, hipStream_t st) {
  if (a.total <= 0) return hipSuccess;
  const int64_t nb = (a.total + block - 1) / block;
  switch (block) {
    case 64: hipLaunchKernelGGL(k_ros3_grid<64>, dim3((unsigned)nb), dim3(64), 0, st, a); break;
    case 128: hipLaunchKernelGGL(k_ros3_grid<128>, dim3((unsigned)nb), dim3(128), 0, st, a); break;
    case 256: hipLaunchKernelGGL(k_ros3_grid<256>, dim3((unsigned)nb), dim3(256), 0, st, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace radm2
