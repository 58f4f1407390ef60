// Workgroup-to-tile mapping for the gather kernels on a multi-XCD GPU.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gkrx {

constexpr uint32_t kXcds = 8;

// Workgroups are dispatched to the XCDs round-robin by id and every XCD has its own L2.
// The gathers of neighbouring tiles touch the same cache lines (the wiring is local), so
// block b, which runs on XCD b % 8, takes tile (b % 8) q + min(b % 8, r) + b / 8 for
// nb = 8 q + r blocks: every XCD works on one contiguous run of tiles.  A bijection of
// [0, nb).
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb)
{
    const uint32_t x = b % kXcds, q = nb / kXcds, r = nb % kXcds;
    return x * q + (x < r ? x : r) + b / kXcds;
}

}  // namespace gkrx
