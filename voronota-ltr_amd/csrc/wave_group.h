// 8- or 16-lane groups of a 64-lane wavefront: the lane-group interface of
// group_clip.h and adjoint.h (lane, ballot, shuffle, shared-memory sync)
// implemented with wavefront intrinsics.
#pragma once

#include <hip/hip_runtime.h>

#include "group_clip.h"

namespace vltr {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int GS>
struct WaveGroupT {
  static constexpr int kSize = GS;
  int base, ln;
  __device__ int lane() const { return ln; }
  __device__ uint32_t ballot(bool p) const {
    return static_cast<uint32_t>((__ballot(p) >> base) & ((1ull << GS) - 1ull));
  }
  __device__ double shfl(double v, int src) const { return __shfl(v, src, GS); }
  __device__ int32_t shfl(int32_t v, int src) const { return __shfl(v, src, GS); }
  __device__ void sync() const { wave_sync(); }
};
using WaveGroup = WaveGroupT<kGroup>;

}  // namespace vltr
