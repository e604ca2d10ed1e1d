// Placement of lane-private working arrays (included before the generated
// mechanism code, which reads and writes every array through RD/WR).
//
// Two placements:
//  * a plain local array (double*): the compiler keeps what it can in VGPRs
//    and AGPRs and spills the rest to scratch;
//  * WsViewT<CPW>: a region of an explicit workspace in HBM, laid out
//    wave-interleaved by cell -- element i of cell c of a wave at
//    wave_base[(region + i) * CPW + c], CPW cells per wave (64 when every
//    lane is a cell; 32 or 16 when a cell spans 2 or 4 lanes, which then
//    read and write the same word) -- and accessed with buffer
//    instructions: a per-wave resource descriptor (SGPRs), the lane's byte
//    offset (one VGPR) and the element offset as a scalar operand, so every
//    access is one instruction moving one contiguous CPW*8-byte segment,
//    with no per-access address arithmetic.  The compiler cannot forward a
//    buffer store to a later load, so the traffic is exactly what the code
//    spells out; launder() additionally ends the reuse window of earlier
//    loads.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace radm2 {

constexpr int WAVE = 64;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
// word 3 of a raw (stride 0) buffer descriptor on CDNA: 32-bit data format
constexpr int kBufferWord3 = 0x00020000;

template <int CPW>
struct WsViewT {
  double* wave_base;            // the wave's workspace (uniform across the wave)
  int region;                   // first element of this array in the workspace
  int lane_off;                 // cell-in-wave * 8 bytes
  int nbytes;                   // size of the wave's workspace in bytes
  __amdgpu_buffer_rsrc_t rsrc;  // descriptor over [wave_base, wave_base + nbytes)
};
using WsView = WsViewT<WAVE>;

template <int CPW = WAVE>
__device__ __forceinline__ WsViewT<CPW> make_view(double* wave_base, int region, int lane,
                                                  int nbytes) {
  WsViewT<CPW> v;
  v.wave_base = wave_base;
  v.region = region;
  v.lane_off = (lane / (WAVE / CPW)) * 8;
  v.nbytes = nbytes;
  v.rsrc = __builtin_amdgcn_make_buffer_rsrc(wave_base, (short)0, nbytes, kBufferWord3);
  return v;
}

template <int CPW>
__device__ __forceinline__ double RD(const WsViewT<CPW>& v, int i) {
  return __builtin_bit_cast(
      double, __builtin_amdgcn_raw_buffer_load_b64(v.rsrc, v.lane_off, (v.region + i) * (CPW * 8), 0));
}
template <int CPW>
__device__ __forceinline__ void WR(const WsViewT<CPW>& v, int i, double x) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), v.rsrc, v.lane_off,
                                        (v.region + i) * (CPW * 8), 0);
}
// The same accesses with the non-temporal cache hint (gfx950's NT bit, bit 1
// of the buffer instructions' cache-policy operand): the line is streamed
// instead of kept, so that it does not push warmer lines out of the L2.
template <int CPW>
__device__ __forceinline__ double RD_stream(const WsViewT<CPW>& v, int i) {
  return __builtin_bit_cast(
      double, __builtin_amdgcn_raw_buffer_load_b64(v.rsrc, v.lane_off, (v.region + i) * (CPW * 8), 2));
}
template <int CPW>
__device__ __forceinline__ void WR_stream(const WsViewT<CPW>& v, int i, double x) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), v.rsrc, v.lane_off,
                                        (v.region + i) * (CPW * 8), 2);
}
// New descriptor from an opaquely "modified" base: loads after this point
// cannot be merged with loads before it.
template <int CPW>
__device__ __forceinline__ void launder(WsViewT<CPW>& v) {
  asm volatile("" : "+s"(v.wave_base));
  v.rsrc = __builtin_amdgcn_make_buffer_rsrc(v.wave_base, (short)0, v.nbytes, kBufferWord3);
}

// A pointer to global memory as a global-address pointer.  The cell arrays'
// pointers come from the domain table, so without the cast the compiler
// would access them as generic (flat) addresses, which it must keep in
// order with every LDS access, and whose loads wait on both memory counters.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}

__device__ __forceinline__ double RD(const double* a, int i) { return a[i]; }
__device__ __forceinline__ void WR(double* a, int i, double x) { a[i] = x; }
__device__ __forceinline__ void launder(double*) {}

// Scheduled reads of the generated code: `const auto x = RDP<i>(a)` in a
// load batch, V(x) at each use.  Here the read happens in the batch; a view
// may overload RDP to return a deferred read that V performs at the use.
template <int I, class A>
__device__ __forceinline__ double RDP(const A& a) {
  return RD(a, I);
}
__device__ __forceinline__ double V(double x) { return x; }

}  // namespace radm2
