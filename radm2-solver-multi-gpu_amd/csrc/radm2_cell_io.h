// Per-lane cell load/store for the RADM2 kernels: domain lookup, input
// gather (SoA, cell fastest), fused rate-constant evaluation, result scatter.
#pragma once
#include "radm2_kernels.h"

namespace radm2 {

// Domain of global work index g: last domain with first <= g.
__device__ __forceinline__ int find_domain(const DomainDesc* __restrict__ doms, int ndom,
                                           int64_t g) {
  int lo = 0, hi = ndom - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (doms[mid].first <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Load cell p of domain d: rate constants (reference RCONST, mgpu.cu:752-919,
// fused here instead of materialised), state vector, fixed species.
template <class KA, class YA>
__device__ __forceinline__ void cell_load(const DomainDesc& d, int64_t p, KA k, YA y,
                                          double& h2o, double& m, CellState& s) {
  const int64_t ld = d.ld;
  m = d.cair[p];
  h2o = d.h2o[p];
  gen::rconst(d.temp[p], m, d.jrate + p, ld, d.rc ? d.rc[p] : d.rc_default, k);
#pragma unroll
  for (int i = 0; i < NVAR; i++) WR(y, i, d.conc[i * ld + p]);
  cell_begin(s, d.opts, d.h0 ? d.h0[p] : 0.0);
}

// Photolysis forcing of cell p of domain d (j1 null when autonomous).
__device__ __forceinline__ Forcing cell_forcing(const DomainDesc& d, int64_t p) {
  Forcing fc;
  fc.j0 = d.jrate + p;
  fc.j1 = d.jrate_end ? d.jrate_end + p : nullptr;
  fc.ld = d.ld;
  fc.dt = d.opts.dt;
  return fc;
}

// writer: the lane that stores (every lane of a cell reads the state, so
// that split entries are broadcast from their owners; one of them writes).
template <class YA>
__device__ __forceinline__ void cell_store(const DomainDesc& d, int64_t p, YA y,
                                           const CellState& s, double hfirst,
                                           unsigned long long* stats, bool writer = true) {
  const int64_t ld = d.ld;
#pragma unroll
  for (int i = 0; i < NVAR; i++) {
    const double v = RD(y, i);
    if (writer) d.conc[i * ld + p] = v;
  }
  if (!writer) return;
  if (d.nstep) d.nstep[p] = s.ns;
  if (d.nacc) d.nacc[p] = s.na;
  if (d.nrej) d.nrej[p] = s.nr;
  if (d.status) d.status[p] = s.status;
  if (d.hexit) d.hexit[p] = s.H;
  if (d.hfirst) d.hfirst[p] = hfirst;
  if (stats) {
    atomicAdd(&stats[0], (unsigned long long)s.ns);
    atomicAdd(&stats[1], (unsigned long long)s.na);
    atomicAdd(&stats[2], (unsigned long long)s.nr);
  }
}

}  // namespace radm2
