// fused.hpp -- FP32 state with gate fusion into dense K-qubit blocks applied
// by FP32 MFMA (the precision pattern).  Shared by fused.hip (kernels +
// C-ABI) and the host-side fusion/matrix code.
#pragma once
#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <vector>

#include "suqa_hip.h"

namespace sqf {

constexpr int kTileLow = 7;        // non-block bits inside a tile: 128 MFMA columns
constexpr int kMinK = 4, kMaxK = 6;

// One fused block: sorted qubits and the 2^K x 2^K complex matrix
// (row-major, acting on the block index sum_j b_j 2^j, b_j = bit qubits[j]).
struct Block {
    uint32_t K;
    uint32_t qubits[kMaxK];
    std::vector<std::complex<double>> m;
    uint32_t ngates;
};

// Greedy fusion with commutation lookahead: a gate joins the open block if
// its qubits fit (at most width in total) and it shares no qubit with any
// earlier gate left out of the block (so moving it forward is exact).
std::vector<Block> fuse(const suqa_hip_gate* g, uint32_t n, uint32_t nq, uint32_t width, uint32_t window);

// 2x2 / controlled matrices of the workload gates (suqa_hip.h SUQA_GATE_*).
void gate_matrix(const suqa_hip_gate& g, std::vector<uint32_t>& qs, std::vector<std::complex<double>>& m);

}  // namespace sqf
