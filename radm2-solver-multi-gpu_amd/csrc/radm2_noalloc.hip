// Device memory of the allocation-free entry point (radm2_integrate_noalloc):
// one file-scope array per device holding the work counter, the domain
// descriptor and the persistent kernel's workspace, so that a call needs no
// hipMalloc.  Each GPU the library runs on holds its copy until the process
// ends.  In host code `scratch` is only a stand-in for the symbol: its
// device address comes from hipGetSymbolAddress, looked up once per device.
#include <mutex>

#include "radm2_kernels.h"

namespace radm2 {

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices];  // device address of `scratch` per device (null: not looked up)

__global__ void k_put_domain(DomainDesc d, DomainDesc* dst) {
  if (threadIdx.x == 0) *dst = d;
}
}  // namespace

hipError_t noalloc_scratch(int device, void** ptr) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch[device] = p;
  }
  *ptr = g_scratch[device];
  return hipSuccess;
}

hipError_t launch_put_domain(const DomainDesc& d, DomainDesc* dst, hipStream_t st) {
  hipLaunchKernelGGL(k_put_domain, dim3(1), dim3(64), 0, st, d, dst);
  return hipGetLastError();
}

}  // namespace radm2
