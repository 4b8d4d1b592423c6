// Device scratch for the allocation-free spread entry point
// (gse_gauss_charge_spread_charge_to_grid_noalloc in capi.cpp).
//
// `scratch` is one file-scope device array: every GPU that loads the library
// holds its own copy of kNoAllocScratchBytes until the process ends.  Host
// code never uses the array's name as an address: the first call on a device
// looks up that device's copy with hipGetSymbolAddress (with the device
// current) and keeps the pointer; every later use goes through that pointer.
#include <hip/hip_runtime.h>

#include <mutex>

#include "gse_internal.h"

namespace gse {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(void** out) {
  *out = nullptr;
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

}  // namespace gse
