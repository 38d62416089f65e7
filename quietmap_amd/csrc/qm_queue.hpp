// Tile queues of the persistent kernels: one contiguous stripe of tiles per XCD (a group's
// blockIdx % 8 is its XCD on MI355X), taken in order by that XCD's groups so that its L2 sees
// neighbouring tiles; a group whose stripe is done steals from the other stripes.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qm {

constexpr int XCDS = 8;

__device__ inline uint32_t next_tile(uint32_t* counters, uint32_t n_tiles, int xcd) {
    for (int k = 0; k < XCDS; ++k) {
        const int q = (xcd + k) % XCDS;
        const uint32_t begin = (uint32_t)(((uint64_t)n_tiles * q) / XCDS);
        const uint32_t end = (uint32_t)(((uint64_t)n_tiles * (q + 1)) / XCDS);
        if (end <= begin) continue;
        if (__hip_atomic_load(&counters[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= end - begin) continue;
        const uint32_t t = atomicAdd(&counters[q], 1u);
        if (t < end - begin) return begin + t;
    }
    return n_tiles;
}

__global__ inline void zero_u32_kernel(uint32_t* p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

}  // namespace qm
