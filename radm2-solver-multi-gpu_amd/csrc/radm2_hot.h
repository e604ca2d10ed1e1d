// Hot set.  Under the LDS placements the most accessed entries of the rate
// constants (array 0), the state (array 1) and the LU (array 2) -- the
// first S of the generated ranking -- live in the lane's slots of the
// wave's LDS area, optionally the next R in a register-resident array of
// the kernel (indexed only by compile-time constants, so it stays in
// VGPRs/AGPRs); every other entry stays in the array's HBM workspace view.
// The lookups are compile-time constants in the unrolled generated code, so
// each RD/WR compiles to one ds_read/ds_write, a register move or one
// buffer access.
//
// LDS layout per wave: slot s of lane l at lds[s * 64 + l] (consecutive lanes
// touch consecutive 8-byte words: one conflict-free b64 access per wave).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "generated/radm2_gen.h"
#include "radm2_ws.h"

namespace radm2 {

typedef __attribute__((address_space(3))) double lds_double;

template <int ARR, int S, int R>
struct HotView {
  WsView ws;        // the array's workspace region (entries in neither tier)
  lds_double* lds;  // this lane's first LDS slot; slot s at lds[s * WAVE]
  double* regs;     // the kernel's register-resident entries (R of them)
};

template <int ARR, int S, int R>
__device__ __forceinline__ HotView<ARR, S, R> make_hot(const WsView& ws, lds_double* wave_lds,
                                                       int lane, double* regs) {
  HotView<ARR, S, R> v;
  v.ws = ws;
  v.lds = wave_lds + lane;
  v.regs = regs;
  return v;
}

template <int ARR, int S, int R>
__device__ __forceinline__ double RD(const HotView<ARR, S, R>& v, int i) {
  const int c = gen::hot_code<S, R>(ARR, i);
  if (c >= 0) return v.lds[c * WAVE];
  if (c <= -2) return v.regs[-2 - c];
  return RD(v.ws, i);
}
template <int ARR, int S, int R>
__device__ __forceinline__ void WR(const HotView<ARR, S, R>& v, int i, double x) {
  const int c = gen::hot_code<S, R>(ARR, i);
  if (c >= 0)
    v.lds[c * WAVE] = x;
  else if (c <= -2)
    v.regs[-2 - c] = x;
  else
    WR(v.ws, i, x);
}
template <int ARR, int S, int R>
__device__ __forceinline__ void launder(HotView<ARR, S, R>& v) {
  launder(v.ws);
}

template <class T>
struct is_ws_like {
  static constexpr bool value = false;
};
template <>
struct is_ws_like<WsView> {
  static constexpr bool value = true;
};
template <int ARR, int S, int R>
struct is_ws_like<HotView<ARR, S, R>> {
  static constexpr bool value = true;
};

}  // namespace radm2
