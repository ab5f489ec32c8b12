// otg_device.hpp -- device helpers shared by the build and query kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace otg {

__device__ __forceinline__ uint32_t bw64(uint64_t x) { return x ? 64u - (uint32_t)__clzll(x) : 0u; }

// si_morton.h:498-574 Encode: bit i of g[d] -> bit i*D + d
template <int D>
__device__ __forceinline__ uint64_t morton_encode(const uint32_t (&g)[D]) {
  uint64_t code = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    uint64_t x = g[d];
    for (int i = 0; i < 32 && (i * D + d) < 64; ++i) code |= ((x >> i) & 1ull) << (i * D + d);
  }
  return code;
}

}  // namespace otg
