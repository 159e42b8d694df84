// deposit.hpp -- scatter of the low bits of x onto the set bits of a mask,
// lowest bit first (the x86 PDEP operation).  Tile kernels use it to turn a
// tile number or a tile element index into state-vector offsets.  The mask is
// uniform across the wave: the loop runs on the scalar unit and, for a per-lane
// x, costs two selects per mask bit.  Kernel arguments are passed as 64-bit
// masks rather than arrays of bit positions: a byte array indexed at run time
// is read with vector loads that drain every memory access in flight.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

__device__ __forceinline__ uint64_t deposit(uint64_t x, uint64_t mask) {
    uint64_t r = 0;
    while (mask) {
        const uint64_t low = mask & (0ull - mask);
        if (x & 1ull) r |= low;
        x >>= 1;
        mask ^= low;
    }
    return r;
}
