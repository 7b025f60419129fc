// The fixed device workspace of vltr_compute_tessellation_device_noalloc:
// one file-scope device array per GPU, from which that entry point carves
// every intermediate buffer of a run, so that it calls no allocator (for
// callers under HIP graph capture, in serving loops, or in harnesses that
// forbid device allocation). The array's name is a host-side stand-in, not a
// device address: its address on each device is taken once, with that device
// current, and only that address is used afterwards.
#include <hip/hip_runtime.h>

#include <atomic>

#include "vltr_internal.h"

namespace vltr {
namespace {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::atomic<void*> g_scratch_ptr[kMaxDevices];

}  // namespace

hipError_t noalloc_scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  if (g_scratch_ptr[device].load() != nullptr) return hipSuccess;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (ptr == nullptr) return hipErrorInvalidSymbol;
  g_scratch_ptr[device].store(ptr);
  return hipSuccess;
}

void* noalloc_scratch(int device) {
  return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device].load() : nullptr;
}

}  // namespace vltr
