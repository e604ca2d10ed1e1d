// Persistent work-stealing kernel: a fixed grid of resident waves; each lane
// (each group of 2 or 4 lanes on the split placements) integrates one cell
// at a time and, when its cell finishes, takes the next cell index from a
// global counter.  Refills are wave-aggregated (one atomic
// per refill round) and happen once at least `refill_min` lanes of the wave
// are idle, trading idle lanes against the cost of the fused rate-constant
// evaluation that a refill runs.  Cells of uneven stiffness (night cells
// finish in a few steps, plume cells take tens) therefore no longer hold a
// whole wave hostage to its slowest lane.
#include "radm2_variant.h"

namespace radm2 {
namespace RADM2_VARIANT_NS {

template <int BLOCK>
__global__ __launch_bounds__(BLOCK, RADM2_WAVES) void k_ros3_persistent(LaunchArgs a) {
  // lanes per cell and cells per wave (one cell per lane except on the
  // split placements)
  constexpr int L = hot_lanes(RADM2_PLACE), CPW = WAVE / L;
#if RADM2_PLACE == RADM2_PLACE_PRIVATE
  double kk[gen::NKSTORE], y[NVAR], lu[LU_NNZ], f[NVAR], k1[NVAR], k2[NVAR], k3[NVAR];
#else
  using VA = WsViewT<CPW>;
  const int64_t wslot = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (BLOCK / WAVE) + threadIdx.x / WAVE));
  constexpr int kWsDoubles = RADM2_METHODS ? WS_DOUBLES_FAMILY : WS_DOUBLES_ROS3;
  constexpr int kWsBytes = kWsDoubles * CPW * 8;
  double* wbase = a.ws + wslot * ((int64_t)kWsDoubles * CPW);
  const int wlane = threadIdx.x & (WAVE - 1);
  VA f = make_view<CPW>(wbase, WS_F, wlane, kWsBytes);
  VA k1 = make_view<CPW>(wbase, WS_K1, wlane, kWsBytes);
  VA k2 = make_view<CPW>(wbase, WS_K2, wlane, kWsBytes);
  VA k3 = make_view<CPW>(wbase, WS_K3, wlane, kWsBytes);
#if RADM2_PLACE >= RADM2_PLACE_LDS160
  // the hot set: kHotSlots doubles per cell in LDS, kHotRegs per lane in
  // registers, kHotSplit per lane split over the cell's lanes
  constexpr int kHotSlots = hot_slots(RADM2_PLACE), kHotRegs = hot_regs(RADM2_PLACE);
  constexpr int kHotSplit = hot_split(RADM2_PLACE);
  __shared__ double hot_lds[kHotSlots * CPW * (BLOCK / WAVE)];
  double hot_rg[kHotRegs > 0 ? kHotRegs : 1];
  double hot_sp[kHotSplit > 0 ? kHotSplit : 1];
  lds_double* wl = (lds_double*)hot_lds + (threadIdx.x / WAVE) * (kHotSlots * CPW);
  auto kk = make_hot<0, kHotSlots, kHotRegs, kHotSplit, L, VA>(
      make_view<CPW>(wbase, WS_K, wlane, kWsBytes), wl, wlane, hot_rg, hot_sp);
  auto y = make_hot<1, kHotSlots, kHotRegs, kHotSplit, L, VA>(
      make_view<CPW>(wbase, WS_Y, wlane, kWsBytes), wl, wlane, hot_rg, hot_sp);
  auto lu = make_hot<2, kHotSlots, kHotRegs, kHotSplit, L, VA>(
      make_view<CPW>(wbase, WS_LU, wlane, kWsBytes), wl, wlane, hot_rg, hot_sp);
#else
  VA kk = make_view<CPW>(wbase, WS_K, wlane, kWsBytes), y = make_view<CPW>(wbase, WS_Y, wlane, kWsBytes);
  VA lu = make_view<CPW>(wbase, WS_LU, wlane, kWsBytes);
#endif
#endif
  double h2o = 0.0, m = 0.0;
  CellState s;
  s.status = ST_OK;
  s.newstep = 0;
  int64_t cell = -1;  // global work index held by this lane's cell, -1 when none
  int di = 0;
  bool exhausted = false;  // wave-uniform: the counter has passed the end
  const int lane = threadIdx.x & 63;
  // the lane of its cell's group that writes the cell's results
  const bool writer = (lane % L) == 0;
  unsigned long long refills = 0;
  for (;;) {
    // every lane of a cell has the same state: idle lanes come in whole
    // groups, so the idle cells number popcount / L and a cell's rank is
    // the idle lanes below its lanes over L
    const bool idle = (s.status != ST_RUNNING);
    const unsigned long long idle_mask = __ballot(idle);
    const int nidle = __popcll(idle_mask) / L;
    if (!exhausted && nidle > 0 && (nidle >= a.refill_min || nidle == CPW)) {
      if (idle && cell >= 0) {
        const DomainDesc& d = a.doms[di];
        cell_store(d, cell - d.first, y, s, s.hfirst, a.stats, writer);
        cell = -1;
      }
      const int leader = __ffsll((long long)idle_mask) - 1;
      unsigned long long base = 0;
      if (lane == leader) base = atomicAdd(a.counter, (unsigned long long)nidle);
      base = __shfl(base, leader);
      const int rank = __popcll(idle_mask & ((1ull << lane) - 1ull)) / L;
      if (idle) {
        const int64_t gi = (int64_t)base + rank;
        if (gi < a.total) {
          di = find_domain(a.doms, a.ndom, gi);
          const DomainDesc& d = a.doms[di];
          cell = gi;
          cell_load(d, gi - d.first, kk, y, h2o, m, s);
        }
      }
      if ((int64_t)(base + (unsigned long long)nidle) >= a.total) exhausted = true;
      refills++;
    }
    if (__ballot(s.status == ST_RUNNING) == 0ull) {
      if (exhausted) break;
      continue;  // all idle but work remains: the next round refills every cell
    }
    if (s.status == ST_RUNNING) {
      const Ros3Opts& o = a.doms[di].opts;
#if RADM2_METHODS
      const Forcing fc = cell_forcing(a.doms[di], cell - a.doms[di].first);
#else
      const Forcing fc{};
#endif
      cell_attempt_top(s, o);
      if (s.status == ST_RUNNING)
        ros3_attempt<RADM2_LROWS, RADM2_METHODS != 0, RADM2_METHODS == 0>(s, o, fc, kk, y, h2o, m, lu, f, k1, k2, k3);
    }
  }
  if (cell >= 0) {
    const DomainDesc& d = a.doms[di];
    cell_store(d, cell - d.first, y, s, s.hfirst, a.stats, writer);
  }
  if (a.stats && lane == 0) atomicAdd(&a.stats[3], refills);
}

}  // namespace RADM2_VARIANT_NS
using namespace RADM2_VARIANT_NS;

template <int BLOCK>
static hipError_t launch_p(int blocks, const LaunchArgs& a, hipStream_t st) {
  hipLaunchKernelGGL(k_ros3_persistent<BLOCK>, dim3(blocks), dim3(BLOCK), 0, st, a);
  return hipGetLastError();
}

hipError_t RADM2_NAME(launch_persistent)(int block, int blocks, const LaunchArgs& a, hipStream_t st) {
  if (a.total <= 0) return hipSuccess;
  switch (block) {
    case 64: return launch_p<64>(blocks, a, st);
#if RADM2_PLACE < RADM2_PLACE_LDS160
    case 256: return launch_p<256>(blocks, a, st);
#endif
    default: return hipErrorInvalidValue;
  }
}

// resident blocks per CU, or minus the HIP error of the occupancy query
int RADM2_NAME(persistent_blocks_per_cu)(int block) {
  int nb = 0;
  hipError_t e = hipErrorInvalidValue;
  switch (block) {
    case 64: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_ros3_persistent<64>, 64, 0); break;
#if RADM2_PLACE < RADM2_PLACE_LDS160
    case 256: e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_ros3_persistent<256>, 256, 0); break;
#endif
    default: break;
  }
  return e == hipSuccess ? nb : -(int)e;
}

}  // namespace radm2
