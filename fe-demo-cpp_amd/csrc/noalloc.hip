// Scratch memory of fem_hip_operator_apply_noalloc: one file-scope device
// array per GPU, so that an apply can run with no device allocation (callers
// under graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address comes from hipGetSymbolAddress, taken once per device with
// that device current, and only that address is ever passed to kernels.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"

namespace femhip {

namespace {

constexpr size_t kNoAllocScratchBytes = size_t{64} << 20;  // 64 MiB per GPU, held until the process ends
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

}  // namespace

size_t noalloc_scratch_capacity() { return kNoAllocScratchBytes; }

// The current device's copy of the array (its address taken on the first call
// on that device).
void* noalloc_scratch_device_ptr() {
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  require(dev >= 0 && dev < kMaxDevices, "apply_noalloc: device id out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    require(p != nullptr && reinterpret_cast<uintptr_t>(p) % 256 == 0, "apply_noalloc: scratch address");
    g_ptr[dev] = p;
  }
  return g_ptr[dev];
}

}  // namespace femhip
