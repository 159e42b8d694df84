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

constexpr int kMinK = 4, kMaxK = 6;

using cd = std::complex<double>;

// One operation before fusion: sorted qubits and its 2^n x 2^n matrix
// (row-major, acting on the local index sum_j b_j 2^j, b_j = bit qs[j]).
struct Op {
    uint32_t n = 0;
    uint32_t qs[kMaxK] = {};
    std::vector<cd> m;
    uint64_t cmask = 0;   // control qubits (all must be 1), outside the matrix
    bool swap = false;    // X on qs[0] with cmask != 0: applied as a swap, never fused
    // a Pauli rotation applied as pair updates, never fused (block passes with
    // sparse rotations): psi'(y) = rc psi(y) + rk (-1)^{|(y ^ rflip) & rpar|} psi(y ^ rflip)
    bool rot = false;
    uint64_t rflip = 0, rpar = 0;
    double rc = 1.0;
    cd rk = 0.0;
};

// One fused block: the same, with exactly K qubits, and how many ops it holds.
struct Block {
    uint32_t K;
    uint32_t qubits[kMaxK];
    std::vector<cd> m;
    uint32_t ngates;
    uint64_t cmask = 0;   // common control of the block's ops: applied where all are 1
    bool swap = false;    // a controlled X (Op::swap) on qubits[0], K = 1, no matrix
    const Op* rot = nullptr;   // a sparse Pauli rotation (Op::rot), no matrix
};

// Greedy fusion with commutation lookahead: an op joins the open block if it
// has the block's control mask, its qubits fit (at most width in total) and
// moving it ahead of every earlier op left out of the block is exact: its
// qubits touch neither the targets nor the controls of those ops, and its
// controls touch none of their targets (ops that share only controls
// commute).
// Blocks are padded to exactly `width` qubits with the lowest free qubits of
// pad_allowed (and never with a control qubit of the block).  Swap ops stay
// single: each becomes a swap block where it is the first op left.
std::vector<Block> fuse(const std::vector<Op>& ops, uint32_t nq, uint32_t width, uint32_t window,
                        uint64_t pad_allowed = ~0ull);

// Matrices of the workload gates (suqa_hip.h SUQA_GATE_*).
Op gate_op(const suqa_hip_gate& g);
// exp(i phase P) of a Pauli string on its non-identity qubits, controlled on
// the gc_mask qubits (Op::cmask); false (and no op) if it is malformed or
// acts on more than kMaxK qubits.
bool pauli_op(const suqa_hip_pauli_op& p, Op& out);
// The same rotation as a sparse op (Op::rot) on the same qubits.
void pauli_rot(const suqa_hip_pauli_op& p, Op& out);
// An op as the pass builder of the block passes sees it: the dense form, and
// the parts a pass may keep outside its tile instead.
//  - ctrl != 0: a multi-controlled X on `tgt`; controls outside the tile
//    become the block's control mask (cx_op over the target and the inside
//    controls).
//  - diag: diag(d0, d1) on qubit dq; outside the tile it is a per-tile scalar
//    (d0 where the bit is 0, d1 where it is 1), which commutes with every op
//    of the pass (no op of a pass targets a qubit outside its tile).
struct PassOp {
    Op op;
    uint64_t ctrl = 0;
    uint32_t tgt = 0;
    bool diag = false;
    uint32_t dq = 0;
    cd d0 = 1.0, d1 = 1.0;
};
PassOp pass_gate(const suqa_hip_gate& g);
// X on tgt controlled by the qubits of `ctrl` (all 1), dense over tgt + ctrl.
Op cx_op(uint64_t ctrl, uint32_t tgt);

// H on q and the controlled phase diag(1,1,1,e^{i phase}) on {a, b}.
Op h_op(uint32_t q);
Op cphase_op(uint32_t a, uint32_t b, double phase);
// The reference's no-swap inverse QFT on qs[0..n) as H / controlled-phase ops.
void qft_inverse_ops(const uint32_t* qs, uint32_t n, std::vector<Op>& out);

}  // namespace sqf
