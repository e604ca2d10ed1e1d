// Kernel-side descriptors and launcher declarations shared by the kernel
// translation units and the C-ABI host code.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

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
  double pad;
  Ros3Opts opts;
};

struct LaunchArgs {
  const DomainDesc* doms;  // device array
  int32_t ndom;
  int32_t refill_min;      // persistent: refill when at least this many lanes idle
  int64_t total;           // sum of domain sizes
  unsigned long long* counter;  // persistent: global work counter (zeroed before launch)
  unsigned long long* stats;    // [4] attempts, accepted, rejected, refills (nullable)
};

// Kernel variants (each in its own translation unit).
hipError_t launch_grid(int block, const LaunchArgs& a, hipStream_t st);
hipError_t launch_persistent(int block, int blocks, const LaunchArgs& a, hipStream_t st);
hipError_t launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t st);
// Occupancy of the persistent kernel for a block size (blocks per CU).
int persistent_blocks_per_cu(int block);

}  // namespace radm2
