// sweep_math.hpp -- per-amplitude arithmetic of the gates, shared by the
// one-gate sweeps (kernels.hip) and the multi-gate tile passes (tiled.hip) so
// that both compute every amplitude with the same operations in the same
// order (bit-identical results; all HIP code is built with -ffp-contract=off).
// Templated on the amplitude type: the FP64 engine instantiates it on double,
// the FP32 engine's gate passes on float (same operations, FP32 rounding).
#pragma once
#include "engine.hpp"

namespace sqe {

constexpr double kTwoSqInv = 0.7071067811865475;   // suqa.cuh:57

// Output of element j of component `im_lane` (0: re, 1: im) whose value is v,
// given the partner k = j ^ X: ps = partner, same component; pc = partner,
// other component.  a_side: j satisfies the gate's selection (the reference's
// index i, the other one is i ^ X).  par = parity of (k & Y|Z positions).
// PC holds c, s (of type T or convertible to it), base_imag and base_neg.
template <int OP, typename T, typename PC>
__device__ __forceinline__ T pair_out_par(bool a_side, bool im_lane, T v, T ps, T pc, uint32_t par, const PC& pc_) {
    if (OP == OP_SWAP) return ps;
    if (OP == OP_Y) {   // suqa_cpu.hpp:116-118
        if (a_side) return im_lane ? pc * (T)-1.0 : pc;   // re_a' = im_b, im_a' = re_b * -1
        return im_lane ? pc : pc * (T)-1.0;               // re_b' = im_a * -1, im_b' = re_a
    }
    if (OP == OP_H) return a_side ? (T)kTwoSqInv * (v + ps) : (T)kTwoSqInv * (ps - v);   // suqa_cpu.hpp:195-198
    // OP_PAULI: psi_j' = c psi_j + base (-1)^{par(k&YZ)} s psi_k
    const T s = (T)pc_.s, c = (T)pc_.c;
    const T t = (pc_.base_neg ^ par) ? -s : s;
    if (!pc_.base_imag) return v * c + ps * t;
    return im_lane ? v * c + pc * t : v * c + pc * (-t);
}

template <int OP>
__device__ __forceinline__ double pair_out(bool a_side, bool im_lane, double v, double ps, double pc, uint64_t k,
                                           const PauliCoef& pc_) {
    const uint32_t par = (OP == OP_PAULI) ? ((uint32_t)__popcll(k & pc_.yz) & 1u) : 0u;
    return pair_out_par<OP>(a_side, im_lane, v, ps, pc, par, pc_);
}

// util_rotate2 (suqa_cpu.hpp:49-54): re' = re*c - im*s, im' = im*c + re*s;
// `other` is the other component of the same amplitude.
template <typename T>
__device__ __forceinline__ T phase_out(bool im_lane, T v, T other, T c, T s) {
    return im_lane ? v * c + other * s : v * c - other * s;
}

}  // namespace sqe
