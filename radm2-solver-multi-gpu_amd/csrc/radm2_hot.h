// Hot set.  Under the LDS placements the most accessed entries of the rate
// constants (array 0), the state (array 1) and the LU (array 2) -- the
// first S of the generated ranking -- live in the cell's slots of the
// wave's LDS area; optionally the next R in a register-resident array of
// the kernel that every lane of the cell holds (indexed only by
// compile-time constants, so it stays in VGPRs/AGPRs); optionally the next
// P*L split over the L lanes that share a cell (see below); every other
// entry stays in the array's HBM workspace view.  The lookups are
// compile-time constants in the unrolled generated code, so each RD/WR
// compiles to one ds_read/ds_write, a register move, a cross-lane move or
// one buffer access.
//
// Cells of L lanes (L = 2 or 4, a group of consecutive lanes inside one
// quad of four): every lane of the group runs the cell's whole computation
// on the same values, so their results are identical bit for bit, but the
// split tier -- register r of group lane q holds split entry r*L + q -- gives
// the cell L times the register capacity of one lane.  A read of a split
// entry broadcasts its owner's register to the group with one DPP quad
// permutation per 32-bit half; a write keeps the new value on the owner
// only (a per-lane select).  All lanes of a group take the same branches,
// so a group is always wholly active or wholly inactive, and the owner of
// a read is active whenever its reader is.
//
// LDS layout per wave: slot s of cell c at lds[s * CPW + c], CPW = 64/L
// cells per wave (the lanes of a cell read the same word: a broadcast;
// consecutive cells touch consecutive 8-byte words).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "generated/radm2_gen.h"
#include "radm2_ws.h"

namespace radm2 {

typedef __attribute__((address_space(3))) double lds_double;

// quad_perm control of a DPP move that gives every lane of an L-lane group
// the value of group lane q (L = 2: lanes {0,1} and {2,3} of each quad)
template <int L>
__device__ __forceinline__ constexpr int group_bcast_ctrl(int q) {
  return L == 4 ? (q | (q << 2) | (q << 4) | (q << 6))
                : (q | (q << 2) | ((q + 2) << 4) | ((q + 2) << 6));
}

template <int L, int Q>
__device__ __forceinline__ double group_bcast_from(double v) {
  const u32x2 w = __builtin_bit_cast(u32x2, v);
  u32x2 r;
  r.x = (unsigned)__builtin_amdgcn_mov_dpp((int)w.x, group_bcast_ctrl<L>(Q), 0xF, 0xF, false);
  r.y = (unsigned)__builtin_amdgcn_mov_dpp((int)w.y, group_bcast_ctrl<L>(Q), 0xF, 0xF, false);
  return __builtin_bit_cast(double, r);
}

// q is a constant once the generated code is inlined; the switch folds away
template <int L>
__device__ __forceinline__ double group_bcast(double v, int q) {
  static_assert(L == 2 || L == 4, "groups of 2 or 4 lanes");
  switch (q) {
    case 0: return group_bcast_from<L, 0>(v);
    case 1: return group_bcast_from<L, 1>(v);
    case 2: return group_bcast_from<L, L == 4 ? 2 : 0>(v);
    default: return group_bcast_from<L, L == 4 ? 3 : 1>(v);
  }
}

template <int ARR, int S, int R, int P = 0, int L = 1, class WS = WsView>
struct HotView {
  WS ws;            // the array's workspace region (entries in no on-chip tier)
  lds_double* lds;  // this cell's first LDS slot; slot s at lds[s * (WAVE / L)]
  double* regs;     // the kernel's replicated register entries (R of them)
  double* split;    // the kernel's split register entries (P per lane)
  int q;            // this lane's index in its cell's group (0 when L = 1)
};

template <int ARR, int S, int R, int P = 0, int L = 1, class WS = WsView>
__device__ __forceinline__ HotView<ARR, S, R, P, L, WS> make_hot(const WS& ws, lds_double* wave_lds,
                                                                 int lane, double* regs,
                                                                 double* split = nullptr) {
  HotView<ARR, S, R, P, L, WS> v;
  v.ws = ws;
  v.lds = wave_lds + lane / L;
  v.regs = regs;
  v.split = split;
  v.q = lane % L;
  return v;
}

// Workspace entries at or beyond rank RADM2_WARM among the workspace's
// entries are accessed with the streaming cache hint, so that the L2 keeps
// the more accessed ones.
#ifndef RADM2_WARM
#define RADM2_WARM (1 << 30)
#endif

template <int ARR, int S, int R, int P, int L, class WS>
__device__ __forceinline__ double RD(const HotView<ARR, S, R, P, L, WS>& v, int i) {
  const int c = gen::hot_code<S, R, P, L>(ARR, i);
  if (c >= 0) return v.lds[c * (WAVE / L)];
  if constexpr (P > 0) {
    if (c <= gen::HOT_SPLIT_BASE) {
      const int j = gen::HOT_SPLIT_BASE - c;
      const double x = group_bcast<L>(v.split[j / L], j % L);
      asm volatile("" : "+v"(v.split[j / L]));  // as in V(SplitRead)
      return x;
    }
  }
  if (c <= gen::HOT_HBM_BASE) {
    if (gen::HOT_HBM_BASE - c >= RADM2_WARM) return RD_stream(v.ws, i);
    return RD(v.ws, i);
  }
  return v.regs[-2 - c];
}
template <int ARR, int S, int R, int P, int L, class WS>
__device__ __forceinline__ void WR(const HotView<ARR, S, R, P, L, WS>& v, int i, double x) {
  const int c = gen::hot_code<S, R, P, L>(ARR, i);
  if (c >= 0) {
    v.lds[c * (WAVE / L)] = x;
    return;
  }
  if constexpr (P > 0) {
    if (c <= gen::HOT_SPLIT_BASE) {
      const int j = gen::HOT_SPLIT_BASE - c;
      double& slot = v.split[j / L];
      slot = v.q == j % L ? x : slot;
      return;
    }
  }
  if (c <= gen::HOT_HBM_BASE) {
    if (gen::HOT_HBM_BASE - c >= RADM2_WARM) WR_stream(v.ws, i, x);
    else WR(v.ws, i, x);
    return;
  }
  v.regs[-2 - c] = x;
}
template <int ARR, int S, int R, int P, int L, class WS>
__device__ __forceinline__ void launder(HotView<ARR, S, R, P, L, WS>& v) {
  launder(v.ws);
}

// Scheduled reads on a hot view (RDP/V of radm2_ws.h): every tier is read
// in the batch except the split register tier, whose cross-lane move is
// deferred to the use: it is a register operation, so there is no latency
// to hide, and a value moved early would hold registers until its use.
template <int L, int J>
struct SplitRead {
  const double* split;
};
template <int I, int ARR, int S, int R, int P, int L, class WS>
__device__ __forceinline__ auto RDP(const HotView<ARR, S, R, P, L, WS>& v) {
  constexpr int c = gen::hot_code<S, R, P, L>(ARR, I);
  if constexpr (P > 0 && c <= gen::HOT_SPLIT_BASE)
    return SplitRead<L, gen::HOT_SPLIT_BASE - c>{v.split};
  else
    return RD(v, I);
}
template <int L, int J>
__device__ __forceinline__ double V(const SplitRead<L, J>& r) {
  double* e = const_cast<double*>(r.split) + J / L;
  const double x = group_bcast_from<L, L == 4 ? J % 4 : J % 2>(*e);
  // the entry comes out of an opaque no-op as a new value, so that the
  // next read's cross-lane move has a different input and is not merged
  // with this one (a merged move would keep the broadcast copy alive)
  asm volatile("" : "+v"(*e));
  return x;
}

template <class T>
struct is_ws_like {
  static constexpr bool value = false;
};
template <int CPW>
struct is_ws_like<WsViewT<CPW>> {
  static constexpr bool value = true;
};
template <int ARR, int S, int R, int P, int L, class WS>
struct is_ws_like<HotView<ARR, S, R, P, L, WS>> {
  static constexpr bool value = true;
};

}  // namespace radm2
