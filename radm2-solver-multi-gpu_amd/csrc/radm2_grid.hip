// One-shot grid kernel: one lane per cell for the whole integration.
// A wave runs until its slowest cell finishes (the reference's schedule,
// mgpu.cu:3504-3641, with the per-cell loop moved onto the lane).  On the
// workspace placements a block borrows one of the workspace's block slots
// for its lifetime (acquire_slot/release_slot), so the workspace is sized by
// the blocks that can be resident, not by the grid.
#include "radm2_variant.h"

// cells spread over several lanes (the split placements) run on the
// persistent schedule only
#if RADM2_PLACE == RADM2_PLACE_SPLIT2 || RADM2_PLACE == RADM2_PLACE_SPLIT4
namespace radm2 {
hipError_t RADM2_NAME(launch_grid)(int, const LaunchArgs&, hipStream_t) { return hipErrorNotSupported; }
int RADM2_NAME(grid_blocks_per_cu)(int) { return -(int)hipErrorNotSupported; }
}  // namespace radm2
#else
namespace radm2 {
namespace RADM2_VARIANT_NS {

// LOCKSTEP: the waves of a block start every attempt together (one barrier
// per attempt), so they walk the attempt's instruction stream side by side
// and share each instruction-cache line; the block runs until its slowest
// cell is done.
template <int BLOCK, bool LOCKSTEP>
__global__ __launch_bounds__(BLOCK, RADM2_WAVES) void k_ros3_grid(LaunchArgs a) {
  const int64_t g = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  const bool valid = g < a.total;
  if (!LOCKSTEP && !valid) return;
  const int64_t gc = valid ? g : a.total - 1;
  const int di = find_domain(a.doms, a.ndom, gc);
  const DomainDesc& d = a.doms[di];
  const int64_t p = gc - d.first;
#if RADM2_PLACE == RADM2_PLACE_PRIVATE
  double kk[gen::NKSTORE], y[NVAR], lu[LU_NNZ], f[NVAR], k1[NVAR], k2[NVAR], k3[NVAR];
#else
  // thread 0 is the block's first cell, valid whenever the block has one;
  // a one-wave block passes the slot on through the wave (no LDS word: the
  // LDS placements' blocks fill their share of the CU's LDS exactly)
  int slot = 0;
  if constexpr (BLOCK == WAVE) {
    if (threadIdx.x == 0) slot = acquire_slot(a.slot_bits, a.slot_words, (int)blockIdx.x);
    slot = __builtin_amdgcn_readfirstlane(slot);
  } else {
    __shared__ int block_slot;
    if (threadIdx.x == 0) block_slot = acquire_slot(a.slot_bits, a.slot_words, (int)blockIdx.x);
    __syncthreads();
    slot = block_slot;
  }
  const int64_t wslot =
      __builtin_amdgcn_readfirstlane((int)((int64_t)slot * (BLOCK / WAVE) + threadIdx.x / WAVE));
  constexpr int kWsDoubles = RADM2_METHODS ? WS_DOUBLES_FAMILY : WS_DOUBLES_ROS3;
  constexpr int kWsBytes = kWsDoubles * WAVE * 8;
  double* wbase = a.ws + wslot * ((int64_t)kWsDoubles * WAVE);
  const int wlane = threadIdx.x & (WAVE - 1);
  WsView f = make_view(wbase, WS_F, wlane, kWsBytes);
  WsView k1 = make_view(wbase, WS_K1, wlane, kWsBytes);
  WsView k2 = make_view(wbase, WS_K2, wlane, kWsBytes);
  WsView k3 = make_view(wbase, WS_K3, wlane, kWsBytes);
#if RADM2_PLACE >= RADM2_PLACE_LDS160
  // the hot set: kHotSlots doubles per lane in LDS, kHotRegs in registers
  constexpr int kHotSlots = hot_slots(RADM2_PLACE), kHotRegs = hot_regs(RADM2_PLACE);
  __shared__ double hot_lds[kHotSlots * BLOCK];
  double hot_rg[kHotRegs > 0 ? kHotRegs : 1];
  lds_double* wl = (lds_double*)hot_lds + (threadIdx.x / WAVE) * (kHotSlots * WAVE);
  auto kk = make_hot<0, kHotSlots, kHotRegs>(make_view(wbase, WS_K, wlane, kWsBytes), wl, wlane, hot_rg);
  auto y = make_hot<1, kHotSlots, kHotRegs>(make_view(wbase, WS_Y, wlane, kWsBytes), wl, wlane, hot_rg);
  auto lu = make_hot<2, kHotSlots, kHotRegs>(make_view(wbase, WS_LU, wlane, kWsBytes), wl, wlane, hot_rg);
#else
  WsView kk = make_view(wbase, WS_K, wlane, kWsBytes), y = make_view(wbase, WS_Y, wlane, kWsBytes);
  WsView lu = make_view(wbase, WS_LU, wlane, kWsBytes);
#endif
#endif
  double h2o, m;
  CellState s;
  cell_load(d, p, kk, y, h2o, m, s);
  if (!valid) s.status = ST_OK;  // padding lane of the last block: never runs
  const Ros3Opts& o = d.opts;
#if RADM2_METHODS
  const Forcing fc = cell_forcing(d, p);
#else
  const Forcing fc{};
#endif
  for (;;) {
    const bool run = s.status == ST_RUNNING;
    if (LOCKSTEP) {
      if (!__syncthreads_or(run)) break;
    } else if (!run) {
      break;
    }
    if (run) {
      cell_attempt_top(s, o);
      if (s.status == ST_RUNNING)
        ros3_attempt<RADM2_LROWS, RADM2_METHODS != 0, RADM2_METHODS == 0>(s, o, fc, kk, y, h2o, m, lu, f, k1, k2, k3);
    }
  }
  if (valid) cell_store(d, p, y, s, s.hfirst, a.stats);
#if RADM2_PLACE != RADM2_PLACE_PRIVATE
  // every wave of the block is done with the slot's workspace
  __threadfence();
  if constexpr (BLOCK != WAVE) __syncthreads();
  if (threadIdx.x == 0) release_slot(a.slot_bits, slot);
#endif
}

}  // namespace RADM2_VARIANT_NS
using namespace RADM2_VARIANT_NS;

// resident blocks per CU, or minus the HIP error of the occupancy query
int RADM2_NAME(grid_blocks_per_cu)(int block) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
  switch (block) {
    case 64: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_ros3_grid<64, false>, 64, 0); break;
#if RADM2_PLACE < RADM2_PLACE_LDS160
    case 256: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_ros3_grid<256, true>, 256, 0); break;
    case 512: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_ros3_grid<512, true>, 512, 0); break;
    case 1024: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_ros3_grid<1024, true>, 1024, 0); break;
#endif
    default: break;
  }
  return e == hipSuccess ? nb : -(int)e;
}

hipError_t RADM2_NAME(launch_grid)(int block, const LaunchArgs& a, hipStream_t st) {
  if (a.total <= 0) return hipSuccess;
  const int64_t nb = (a.total + block - 1) / block;
  switch (block) {
    case 64: hipLaunchKernelGGL((k_ros3_grid<64, false>), dim3((unsigned)nb), dim3(64), 0, st, a); break;
#if RADM2_PLACE < RADM2_PLACE_LDS160
    case 256: hipLaunchKernelGGL((k_ros3_grid<256, true>), dim3((unsigned)nb), dim3(256), 0, st, a); break;
#endif
#if RADM2_PLACE < RADM2_PLACE_LDS160
    case 512: hipLaunchKernelGGL((k_ros3_grid<512, true>), dim3((unsigned)nb), dim3(512), 0, st, a); break;
#endif
#if RADM2_PLACE < RADM2_PLACE_LDS160
    case 1024: hipLaunchKernelGGL((k_ros3_grid<1024, true>), dim3((unsigned)nb), dim3(1024), 0, st, a); break;
#endif
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace radm2
#endif
