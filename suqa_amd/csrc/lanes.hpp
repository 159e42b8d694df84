// lanes.hpp -- the wavefront butterfly sum without the LDS path of
// ds_bpermute: DPP moves inside a row of 16 lanes, the CDNA4 permlane swaps
// across rows.  In lane_xor<O>, lane l receives the value of lane l ^ O.  The moves'
// lane layouts (quad_perm, row_shl/row_shr with a bank mask, row_ror:8,
// permlane16/32_swap) were checked on gfx950 by writing the lane index
// through each of them; tests/test_gpu_reduction_order.py pins the sums
// wave_sum64 forms, bit for bit, to a host restatement of the butterfly.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sqe {

__device__ __forceinline__ uint32_t dlo(double v) { return (uint32_t)(uint64_t)__double_as_longlong(v); }
__device__ __forceinline__ uint32_t dhi(double v) { return (uint32_t)((uint64_t)__double_as_longlong(v) >> 32); }
__device__ __forceinline__ double dmake(uint32_t lo, uint32_t hi) {
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// one dword of lane l ^ O, O = 1, 2, 4 or 8 (DPP inside a row of 16 lanes)
template <int O>
__device__ __forceinline__ uint32_t row_xor(uint32_t v) {
    static_assert(O == 1 || O == 2 || O == 4 || O == 8, "row_xor: 1, 2, 4 or 8");
    if constexpr (O == 1) {
        return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    } else if constexpr (O == 2) {
        return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    } else if constexpr (O == 4) {
        // banks 0 and 2 of a row (lanes with bit 2 clear) read lane + 4 (row_shl:4),
        // banks 1 and 3 read lane - 4 (row_shr:4)
        const int t = __builtin_amdgcn_update_dpp((int)v, (int)v, 0x104, 0xF, 0x5, false);
        return (uint32_t)__builtin_amdgcn_update_dpp(t, (int)v, 0x114, 0xF, 0xA, false);
    } else {
        return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x128, 0xF, 0xF, false);  // row_ror:8
    }
}

// the value of lane l ^ O for a double, O = 1, 2, 4 or 8
template <int O>
__device__ __forceinline__ double lane_xor(double v) {
    return dmake(row_xor<O>(dlo(v)), row_xor<O>(dhi(v)));
}

// v + (lane l ^ o's v) for o = 32, 16, ..., 1: every lane ends with the same
// sum of all 64, formed in the order of the __shfl_xor butterfly of that
// order (each step adds two values, and a + b == b + a).
__device__ __forceinline__ double wave_sum64(double v) {
    {
        const auto l = __builtin_amdgcn_permlane32_swap(dlo(v), dlo(v), false, false);
        const auto h = __builtin_amdgcn_permlane32_swap(dhi(v), dhi(v), false, false);
        v = dmake(l[0], h[0]) + dmake(l[1], h[1]);
    }
    {
        const auto l = __builtin_amdgcn_permlane16_swap(dlo(v), dlo(v), false, false);
        const auto h = __builtin_amdgcn_permlane16_swap(dhi(v), dhi(v), false, false);
        v = dmake(l[0], h[0]) + dmake(l[1], h[1]);
    }
    v += lane_xor<8>(v);
    v += lane_xor<4>(v);
    v += lane_xor<2>(v);
    v += lane_xor<1>(v);
    return v;
}

}  // namespace sqe
