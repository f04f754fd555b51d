// cent_noalloc.hip - the fixed device scratch of dasp_cent_run_device_noalloc.
//
// One file-scope device array per GPU, sized kNoAllocScratchBytes, holds the scratch of the
// allocation-free device run (images handed to the correlation stage, correlation images, the
// resampled phase grid, the general path's hll), for callers that cannot allocate device memory.
// In host code the array's name is not a device address: its address on each device is looked
// up once (hipGetSymbolAddress, with that device current) and only that address is used.
#include <hip/hip_runtime.h>

#include <mutex>

#include "cent_internal.h"

namespace dasp {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};
}  // namespace

// Address of the scratch array on the current device, looked up on the first call for that device.
hipError_t noalloc_scratch(void **out) {
  *out = nullptr;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return hipSuccess;
}

}  // namespace dasp
