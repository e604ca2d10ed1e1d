// Kernel-side descriptors and launcher declarations shared by the kernel
// translation units and the C-ABI host code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "radm2.h"
#include "radm2_solver.h"

namespace radm2 {

// One integration domain: SoA arrays with the cell index fastest,
// x[s*ld + p], all device pointers.  `first` is the domain's offset in the
// launch's global work index space (prefix sum of sizes).
struct DomainDesc {
  int64_t first, n, ld;
  double* conc;          // [59][ld] in -> out
  const double* temp;    // [ld] K
  const double* cair;    // [ld] molec/cm3
  const double* h2o;     // [ld] molec/cm3
  const double* jrate;   // [21][ld] s-1, reference jv order
  const double* rc;      // [ld] or null -> rc_default
  const double* h0;      // [ld] or null: proposed initial step per cell
  int32_t* nstep;        // outputs, each nullable
  int32_t* nacc;
  int32_t* nrej;
  int32_t* status;
  double* hexit;
  double* hfirst;        // proposal after the first accepted step (learning target)
  double rc_default;
  const double* jrate_end;  // [21][ld] J at t = dt (non-autonomous) or null
  Ros3Opts opts;
};

struct LaunchArgs {
  const DomainDesc* doms;  // device array
  int32_t ndom;
  int32_t refill_min;      // persistent: refill when at least this many lanes idle
  int64_t total;           // sum of domain sizes
  unsigned long long* counter;  // persistent: global work counter (zeroed before launch)
  unsigned long long* stats;    // [4] attempts, accepted, rejected, refills (nullable)
  double* ws;                   // explicit workspace (WS placement), ws_doubles per lane
  int64_t ws_waves;             // wave slots the workspace holds
  // grid schedule on a workspace placement: occupancy bitmap of the
  // workspace's block slots (BLOCK/64 wave slots each; bit set = taken),
  // slot_words 64-bit words, bits past the last slot preset
  unsigned long long* slot_bits;
  int32_t slot_words;
};

// Explicit-workspace layout per cell (doubles): the stored rate-constant
// values (gen::NKSTORE of them: the literal rate constants are written into
// the code, those of the form A*exp(c/T) share their exponential), state,
// LU, four stage vectors.
constexpr int WS_K = 0;
constexpr int WS_Y = WS_K + gen::NKSTORE;
constexpr int WS_LU = WS_Y + NVAR;
constexpr int WS_F = WS_LU + LU_NNZ;
constexpr int WS_K1 = WS_F + NVAR;
constexpr int WS_K2 = WS_K1 + NVAR;
constexpr int WS_K3 = WS_K2 + NVAR;
// doubles per lane: the Ros3 kernels hold K1..K3; the kernels of KPP's
// whole Rosenbrock family six stage vectors (Rodas4), stage s at
// WS_K1 + s*NVAR, and dF/dT at WS_K1 + 6*NVAR
constexpr int WS_DOUBLES_ROS3 = WS_K3 + NVAR;
constexpr int WS_DOUBLES_FAMILY = WS_K1 + 7 * NVAR;

// Placement of the lane-private working arrays
// (preprocessor tests in the kernels use the RADM2_PLACE_* macros of radm2.h)
constexpr int PLACE_PRIVATE = RADM2_PLACE_PRIVATE;  // compiler-managed (VGPR/AGPR, scratch spills)
constexpr int PLACE_WS = RADM2_PLACE_WORKSPACE;     // explicit wave-interleaved HBM workspace
// workspace with the most accessed entries in LDS (64-lane blocks only):
// 160 doubles per lane (two blocks per CU) or 312 (one block per CU)
constexpr int PLACE_HOT160 = RADM2_PLACE_LDS160;
constexpr int PLACE_HOT312 = RADM2_PLACE_LDS312;
// Hot-set configuration (LDS slots per cell, replicated register entries
// per lane, split register entries per lane, lanes per cell) of each
// placement; measured at 1M cells with one cell per lane: more LDS per
// lane (fewer resident waves) or more registers (spills) or fewer LDS
// slots (more waves, more HBM traffic) are all slower than (160, 112), which
// leaves the Ros3 kernels the registers for load batches of 20.
// The split placements spread a cell over 2 or 4 lanes so that four waves
// per CU (one per SIMD) hold as many cells (2 lanes) or half as many (4
// lanes) as the one-lane placements' two.
constexpr int PLACE_HOT160R112 = RADM2_PLACE_LDS160_REG112;
constexpr int PLACE_SPLIT2 = RADM2_PLACE_SPLIT2;
constexpr int PLACE_SPLIT4 = RADM2_PLACE_SPLIT4;
constexpr int kHotConfig[7][4] = {{0, 0, 0, 1},   {0, 0, 0, 1},     {160, 0, 0, 1},
                                  {312, 0, 0, 1}, {160, 112, 0, 1}, {160, 0, 128, 2},
                                  {320, 0, 128, 4}};
constexpr int PLACE_HOT_LAST = PLACE_SPLIT4;
constexpr int hot_slots(int placement) { return kHotConfig[placement][0]; }
constexpr int hot_regs(int placement) { return kHotConfig[placement][1]; }
constexpr int hot_split(int placement) { return kHotConfig[placement][2]; }
constexpr int hot_lanes(int placement) { return kHotConfig[placement][3]; }

// Kernel variants: one translation unit per (placement P, register budget
// W = minimum resident waves per SIMD (1, 2, 4, 8), launder interval R of the
// compiler-scheduled kernels, 0 in every built variant); block sizes 64 and
// 256 in each.  The Ros3 variants compile Ros3 alone; the family variant
// (suffix _m1) runs every method of KPP's Rosenbrock family from one table
// of coefficients, at the cost of registers and stage storage that would
// slow Ros3.
#define RADM2_DECLARE_VARIANT_NAMED(S)                                                   \
  hipError_t launch_grid##S(int block, const LaunchArgs& a, hipStream_t st);             \
  int grid_blocks_per_cu##S(int block);                                                  \
  hipError_t launch_persistent##S(int block, int blocks, const LaunchArgs& a, hipStream_t st); \
  int persistent_blocks_per_cu##S(int block);
#define RADM2_DECLARE_VARIANT(P, W, R) RADM2_DECLARE_VARIANT_NAMED(_p##P##_w##W##_r##R)
RADM2_DECLARE_VARIANT(0, 1, 0)
RADM2_DECLARE_VARIANT(0, 2, 0)
RADM2_DECLARE_VARIANT(0, 4, 0)
RADM2_DECLARE_VARIANT(0, 8, 0)
RADM2_DECLARE_VARIANT(1, 1, 0)
RADM2_DECLARE_VARIANT(1, 2, 0)
RADM2_DECLARE_VARIANT(1, 4, 0)
RADM2_DECLARE_VARIANT(2, 1, 0)
RADM2_DECLARE_VARIANT(3, 1, 0)
RADM2_DECLARE_VARIANT(4, 1, 0)
RADM2_DECLARE_VARIANT(5, 1, 0)
RADM2_DECLARE_VARIANT(6, 1, 0)
RADM2_DECLARE_VARIANT_NAMED(_p4_w1_r0_m1)
#undef RADM2_DECLARE_VARIANT
#undef RADM2_DECLARE_VARIANT_NAMED

struct KernelVariant {
  int placement;
  int waves;
  int launder_rows;
  int family;      // 1: KPP's five methods; 0: Ros3 only
  int ws_doubles;  // workspace doubles per cell
  int lanes;       // lanes per cell
  hipError_t (*grid)(int, const LaunchArgs&, hipStream_t);
  hipError_t (*persistent)(int, int, const LaunchArgs&, hipStream_t);
  int (*blocks_per_cu)(int);
  int (*grid_blocks_per_cu)(int);
};
constexpr int kNumVariants = 13;
// workspace bytes of one wave of a variant
inline size_t ws_wave_bytes(const KernelVariant* v) {
  return (size_t)v->ws_doubles * (WAVE / v->lanes) * sizeof(double);
}
const KernelVariant* kernel_variants();

hipError_t launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t st);
// clear an occupancy bitmap of `nslots` slots over `nwords` words (the bits
// past the last slot set)
hipError_t launch_slot_init(unsigned long long* bits, int nwords, int64_t nslots, hipStream_t st);

// Block-slot pool of the grid schedule: a block takes a free slot of the
// bitmap when it starts and returns it when its waves are done with the
// slot's workspace, so the workspace holds as many block slots as can be
// resident at once instead of one per block of the grid.
__device__ __forceinline__ int acquire_slot(unsigned long long* bits, int nwords, int start) {
  for (int w = start % nwords;; w = (w + 1 == nwords) ? 0 : w + 1) {
    unsigned long long cur = __hip_atomic_load(bits + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (cur != ~0ull) {
      const int b = __ffsll((long long)~cur) - 1;
      const unsigned long long bit = 1ull << b;
      const unsigned long long old = atomicOr(bits + w, bit);
      if (!(old & bit)) return w * 64 + b;
      cur = old | bit;
    }
  }
}
// The atomic's returned value is consumed, so the wave waits for the
// release to complete before it moves on (and exits).
__device__ __forceinline__ void release_slot(unsigned long long* bits, int slot) {
  const unsigned long long old = atomicAnd(bits + (slot >> 6), ~(1ull << (slot & 63)));
  asm volatile("" ::"v"(old));
}

hipError_t measure_fp64_peak(hipStream_t st, int cus, double* fp64_tflops);

// The allocation-free entry point's device array (radm2_noalloc.hip): the
// work counter at byte 0, the domain descriptor at kNoallocDomOffset, the
// workspace from kNoallocWsOffset to the end.
constexpr size_t kNoallocScratchBytes = size_t(128) << 20;
constexpr size_t kNoallocDomOffset = 256;
constexpr size_t kNoallocWsOffset = 4096;
static_assert(kNoallocDomOffset + sizeof(DomainDesc) <= kNoallocWsOffset, "scratch header");
// This device's address of the array (looked up once per device, with the
// device current).
hipError_t noalloc_scratch(int device, void** ptr);
// Writes `d` to device memory `dst` with a kernel (no copy from host memory).
hipError_t launch_put_domain(const DomainDesc& d, DomainDesc* dst, hipStream_t st);

}  // namespace radm2
