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
// fused here instead of materialised), state vector, fixed species.  A
// workspace store may alias global memory as far as the compiler knows, so
// a load after one waits for nothing but cannot be issued before it: the
// descriptor fields and scalars are read before the first store and the
// concentrations 16 at a time ahead of their stores, a few memory round
// trips per cell instead of one per concentration.
template <class KA, class YA>
__device__ __forceinline__ void cell_load(const DomainDesc& d, int64_t p, KA k, YA y,
                                          double& h2o, double& m, CellState& s) {
  const int64_t ld = d.ld;
  const auto conc = gmem(d.conc);
  const auto rcp = gmem(d.rc);
  const auto h0p = gmem(d.h0);
  const Ros3Opts o = d.opts;
  const double rc_default = d.rc_default;
  const double T = gmem(d.temp)[p];
  m = gmem(d.cair)[p];
  h2o = gmem(d.h2o)[p];
  const double rc = rcp ? rcp[p] : rc_default;
  const double h0 = h0p ? h0p[p] : 0.0;
  // through a global-address pointer, so that its loads are global ones
  gen::rconst(T, m, (const double*)(gmem(d.jrate) + p), ld, rc, k);
  constexpr int CH = 16;
#pragma unroll
  for (int b = 0; b < NVAR; b += CH) {
    double c[CH];
#pragma unroll
    for (int j = 0; j < CH; j++)
      if (b + j < NVAR) c[j] = conc[(b + j) * ld + p];
#pragma unroll
    for (int j = 0; j < CH; j++)
      if (b + j < NVAR) WR(y, b + j, c[j]);
  }
  cell_begin(s, o, h0);
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
// The state is read 16 entries at a time ahead of their stores, and the
// descriptor's pointers before the first store (as in cell_load).
template <class YA>
__device__ __forceinline__ void cell_store(const DomainDesc& d, int64_t p, YA y,
                                           const CellState& s, double hfirst,
                                           unsigned long long* stats, bool writer = true) {
  const int64_t ld = d.ld;
  const auto conc = gmem(d.conc);
  const auto nstep = gmem(d.nstep);
  const auto nacc = gmem(d.nacc);
  const auto nrej = gmem(d.nrej);
  const auto status = gmem(d.status);
  const auto hexit = gmem(d.hexit);
  const auto hf = gmem(d.hfirst);
  constexpr int CH = 16;
#pragma unroll
  for (int b = 0; b < NVAR; b += CH) {
    double v[CH];
#pragma unroll
    for (int j = 0; j < CH; j++)
      if (b + j < NVAR) v[j] = RD(y, b + j);
#pragma unroll
    for (int j = 0; j < CH; j++)
      if (b + j < NVAR && writer) conc[(b + j) * ld + p] = v[j];
  }
  if (!writer) return;
  if (nstep) nstep[p] = s.ns;
  if (nacc) nacc[p] = s.na;
  if (nrej) nrej[p] = s.nr;
  if (status) status[p] = s.status;
  if (hexit) hexit[p] = s.H;
  if (hf) hf[p] = hfirst;
  if (stats) {
    atomicAdd(&stats[0], (unsigned long long)s.ns);
    atomicAdd(&stats[1], (unsigned long long)s.na);
    atomicAdd(&stats[2], (unsigned long long)s.nr);
  }
}

}  // namespace radm2
