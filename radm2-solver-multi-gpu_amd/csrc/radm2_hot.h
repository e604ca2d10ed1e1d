// LDS-resident hot set.  Under this placement the most accessed entries of
// the rate constants (array 0), the state (array 1) and the LU (array 2) --
// the first S of the generated ranking -- live in the lane's slots of the
// wave's LDS area; every other entry stays in the array's HBM workspace
// view.  Slot lookups are compile-time constants in the unrolled generated
// code, so each RD/WR compiles to one ds_read/ds_write or one buffer access.
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

template <int ARR, int S>
struct HotView {
  WsView ws;        // the array's workspace region (entries without a slot)
  lds_double* lds;  // this lane's first slot; slot s at lds[s * WAVE]
};

template <int ARR, int S>
__device__ __forceinline__ HotView<ARR, S> make_hot(const WsView& ws, lds_double* wave_lds, int lane) {
  HotView<ARR, S> v;
  v.ws = ws;
  v.lds = wave_lds + lane;
  return v;
}

template <int ARR, int S>
__device__ __forceinline__ double RD(const HotView<ARR, S>& v, int i) {
  const int s = gen::hot_slot<S>(ARR, i);
  if (s >= 0) return v.lds[s * WAVE];
  return RD(v.ws, i);
}
template <int ARR, int S>
__device__ __forceinline__ void WR(const HotView<ARR, S>& v, int i, double x) {
  const int s = gen::hot_slot<S>(ARR, i);
  if (s >= 0)
    v.lds[s * WAVE] = x;
  else
    WR(v.ws, i, x);
}
template <int ARR, int S>
__device__ __forceinline__ void launder(HotView<ARR, S>& v) {
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
template <int ARR, int S>
struct is_ws_like<HotView<ARR, S>> {
  static constexpr bool value = true;
};

}  // namespace radm2
