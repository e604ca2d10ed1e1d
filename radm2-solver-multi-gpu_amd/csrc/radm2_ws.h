// Placement of lane-private working arrays (included before the generated
// mechanism code, which reads and writes every array through RD/WR).
//
// Two placements:
//  * a plain local array (double*): the compiler keeps what it can in VGPRs
//    and AGPRs and spills the rest to scratch;
//  * WsView: a region of an explicit workspace in HBM, laid out
//    wave-interleaved -- element i of lane l of a wave at
//    wave_base[(region + i) * 64 + l] -- and accessed with buffer
//    instructions: a per-wave resource descriptor (SGPRs), the lane's byte
//    offset (one VGPR) and the element offset as a scalar operand, so every
//    access is one instruction moving one contiguous 512-byte segment, with
//    no per-access address arithmetic.  The compiler cannot forward a buffer
//    store to a later load, so the traffic is exactly what the code spells
//    out; launder() additionally ends the reuse window of earlier loads.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace radm2 {

constexpr int WAVE = 64;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
// word 3 of a raw (stride 0) buffer descriptor on CDNA: 32-bit data format
constexpr int kBufferWord3 = 0x00020000;

struct WsView {
  double* wave_base;            // the wave's workspace (uniform across the wave)
  int region;                   // first element of this array in the workspace
  int lane_off;                 // lane * 8 bytes
  int nbytes;                   // size of the wave's workspace in bytes
  __amdgpu_buffer_rsrc_t rsrc;  // descriptor over [wave_base, wave_base + nbytes)
};

__device__ __forceinline__ WsView make_view(double* wave_base, int region, int lane, int nbytes) {
  WsView v;
  v.wave_base = wave_base;
  v.region = region;
  v.lane_off = lane * 8;
  v.nbytes = nbytes;
  v.rsrc = __builtin_amdgcn_make_buffer_rsrc(wave_base, (short)0, nbytes, kBufferWord3);
  return v;
}

__device__ __forceinline__ double RD(const WsView& v, int i) {
  return __builtin_bit_cast(
      double, __builtin_amdgcn_raw_buffer_load_b64(v.rsrc, v.lane_off, (v.region + i) * (WAVE * 8), 0));
}
__device__ __forceinline__ void WR(const WsView& v, int i, double x) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, x), v.rsrc, v.lane_off,
                                        (v.region + i) * (WAVE * 8), 0);
}
// New descriptor from an opaquely "modified" base: loads after this point
// cannot be merged with loads before it.
__device__ __forceinline__ void launder(WsView& v) {
  asm volatile("" : "+s"(v.wave_base));
  v.rsrc = __builtin_amdgcn_make_buffer_rsrc(v.wave_base, (short)0, v.nbytes, kBufferWord3);
}

__device__ __forceinline__ double RD(const double* a, int i) { return a[i]; }
__device__ __forceinline__ void WR(double* a, int i, double x) { a[i] = x; }
__device__ __forceinline__ void launder(double*) {}

}  // namespace radm2
