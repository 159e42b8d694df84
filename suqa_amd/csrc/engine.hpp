// engine.hpp -- internals of the gfx950 state-vector engine shared by the
// kernels (kernels.hip) and the C-ABI (capi.cpp).
//
// Every sweep visits a *selection*: the basis indices i with (i & M) == V,
// enumerated densely by inserting the fixed bits of M into a counter k.
// This is the reference's predicate form (e.g. suqa_cpu.hpp:86-92), but the
// GPU only touches the selected amplitudes instead of testing all N.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "suqa_hip.h"

namespace sqe {

constexpr int kBlock = 256;          // 4 waves of 64 lanes
constexpr uint64_t kPartChunk = 2048; // partial pairs summed per block in the reduction tree

// Sorted fixed bit positions of a selection.
struct Fixed {
    uint32_t n;
    uint32_t pos[32];
};

struct Sel {
    Fixed fb;
    uint64_t fval;     // values of the fixed bits
    uint64_t count;    // number of selected indices = N >> fb.n
};

// Selection {i < 2^nq : (i & M) == V}; returns false when it is empty.
bool make_sel(uint32_t nq, uint64_t M, uint64_t V, Sel* out);

enum PairOp : int { OP_SWAP = 0, OP_Y = 1, OP_H = 2, OP_PAULI = 3 };

// exp(i phi P) on a pair (a, b = a ^ xy):
//   psi'_a = c psi_a + zeta_a s psi_b, zeta_a = base (-1)^{par(b & yz)},
//   base = i^{nY+1}: base_imag = (nY even), base_neg = ((nY+1) mod 4 >= 2).
struct PauliCoef {
    double c, s;
    uint64_t yz;       // Y|Z positions
    uint32_t ny_odd;   // nY mod 2 == par(xy & yz)
    uint32_t base_imag;
    uint32_t base_neg;
};

// A gate sweep as the launchers receive it: a pair operation (OP_*) on the
// pairs (i, i ^ X) with (i & M) == V, or a phase rotation (kPhaseOp) of the
// amplitudes with (i & M) == V, s negated where par(i & zmask) == 1.
constexpr int kPhaseOp = 4;
// Queued operations other than gates, for states of one small tile only
// (tiny_prog below): a measurement of qubit q with draw c, its outcome into
// outcome slot `slot`, and for a reset X on q over the gc bits (M) when the
// outcome is 1; qms::apply_W on the acceptance pairs.
constexpr int kMeasOp = 5, kWOp = 6;
constexpr uint32_t kNoRef = 0xffffffffu;
struct GateDesc {
    int kind;
    uint64_t M, V, X;
    PauliCoef pc;        // pair operations
    double c, s;         // phase rotations; kMeasOp: c = the draw
    uint64_t zmask;
    const char* tag;
    double alg_bytes;
    // tiny_prog operations: a gate with cond_ref != kNoRef applies only when
    // outcome slot cond_ref reads 1; kMeasOp: q, slot, reset; kWOp: the
    // (fs1, fs2) table, the register masks and E_old (a value, the outcome
    // slots eold_ref .. + eold_bits - 1 read LSB first, or omask bits)
    uint32_t cond_ref = kNoRef;
    uint32_t q = 0, slot = 0, reset = 0;
    uint32_t meas_group = 0;   // kMeasOp: on the first of a call's measurements, their number
    const double* tab = nullptr;
    uint32_t wmask = 0, b0 = 0, omask = 0, ob0 = 0, L = 0;
    uint32_t eold = 0, eold_ref = kNoRef, eold_bits = 0;
};

struct KStat {
    std::string name;
    uint64_t launches = 0;
    double total_ms = 0.0;
    double alg_bytes = 0.0;
};

}  // namespace sqe

struct suqa_hip_ctx {
    int device = 0;
    int num_cu = 256;
    hipStream_t stream = nullptr;
    uint32_t nq = 0;
    uint64_t N = 0;
    double* d_data = nullptr;
    double* d_re = nullptr;
    double* d_im = nullptr;
    // reductions: per-block partials (2 per block), final result, host copy
    double* d_partials = nullptr;      // 2 * sqe::partial_pairs_capacity(N) doubles
    double* d_result = nullptr;
    double* h_result = nullptr;     // pinned, 4 doubles
    uint32_t* d_flags = nullptr;    // outcomes of launch_measure_small (sqe::kMeasMax)
    uint32_t* h_flags = nullptr;    // pinned copy
    // outcome slots of queued measurements (tiny_prog states): a ring of
    // sqe::kSlots outcomes on the device, a pinned copy, the next free slot
    uint32_t* d_slots = nullptr;
    uint32_t* h_slots = nullptr;
    uint32_t slot_next = 0;
    // qms::apply_W coefficient table (fs1, fs2) per (E_old, E_new)
    double* d_wtab = nullptr;
    uint64_t wtab_L = 0;
    bool wtab_valid = false;
    double wtab_c = 0.0;
    uint32_t wtab_mask = 0, wtab_b0 = 0;
    double* d_wrow = nullptr;       // slow path: one row, L > 256
    uint64_t wrow_cap = 0;
    // one-shot QFT phase table (suqa.cu:611-625)
    bool rphase_filled = false;
    std::vector<double> rphase;
    // gate passes: with tile_bits > 0, pair and phase sweeps are queued and
    // applied in passes of up to kMaxPass gates over tiles of 2^tile_bits
    // amplitudes held in LDS (tiled.hip); every other operation flushes first
    uint32_t tile_bits = 0;
    bool pauli_reorder = true;   // run_paulis packs passes with commuting ops moved ahead
    // span tiles: a pass's tile coordinates above the low qubits are the GF(2)
    // span of its gates' flip masks (a weight-w Pauli string takes one
    // coordinate, not w); off: one coordinate per flipped qubit
    bool span_tiles = true;
    // the FP32 engine's gate queue (sqe::f32_gate_queue): no FP64 state of its
    // own; gates queue as above and its passes run on the float state f32_re,
    // f32_im (one-gate sweeps and every other operation are refused)
    bool f32 = false;
    float* f32_re = nullptr;
    float* f32_im = nullptr;
    // flush packs passes with commuting gates moved ahead (the same product,
    // another rounding): FP32 queues unless SUQA_HIP_GATE_ORDER=0, FP64
    // contexts when suqa_hip_set_gate_order asks for it (the suqa:: API never does)
    bool gate_reorder = false;
    std::vector<sqe::GateDesc> queue;
    uint64_t passes = 0, pass_gates = 0, single_sweeps = 0;
    // captured launch sequences (suqa_hip_graph_*): replayable graphs on the
    // context stream; they hold the state's device addresses
    bool capturing = false;
    std::vector<hipGraphExec_t> graphs;
    // timing
    bool timing = false;
    struct Rec {
        int stat;
        hipEvent_t e0, e1;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> evpool;
    std::vector<sqe::KStat> stats;
};

namespace sqe {

// The selection of suqa::prob_filter (suqa.cu:706-713): bits qs[k] of the
// index must read q_mask[k].  False when a qubit index is 32 or more.
inline bool prob_filter_masks(const uint32_t* qs, const uint32_t* q_mask, uint32_t n, uint32_t* mqs,
                              uint32_t* mask) {
    uint32_t m = 0u, v = 0u;
    for (uint32_t k = 0; k < n; ++k) {
        if (qs[k] >= 32) return false;
        m |= 1u << qs[k];
    }
    for (uint32_t k = 0; k < n; ++k)
        if (q_mask[k]) v |= q_mask[k] << qs[k];
    *mqs = m;
    *mask = v;
    return true;
}

// Launchers (kernels.hip).  All enqueue on ctx->stream and return a HIP error.
hipError_t launch_init(suqa_hip_ctx* c);
hipError_t launch_scale(suqa_hip_ctx* c, double f);
hipError_t launch_pair(suqa_hip_ctx* c, int op, const Sel& sel, uint64_t xmask, const PauliCoef& pc,
                       const char* tag, double alg_bytes);
hipError_t launch_phase(suqa_hip_ctx* c, const Sel& sel, double cs, double sn, uint64_t zmask,
                        const char* tag, double alg_bytes);
hipError_t launch_phase_list(suqa_hip_ctx* c, const Sel& sel, uint32_t q0, uint32_t size_mask,
                             const double* cs, const double* sn);
// sum |psi|^2 over sel into d_result[0] (split_bit == 0) or into
// d_result[0] (bit clear) and d_result[1] (bit set).
hipError_t launch_norm(suqa_hip_ctx* c, const Sel& sel, uint64_t split_bit, const char* tag,
                       double alg_bytes);
// The reduction of launch_norm on any state: k_norm writes one (sum, sum) pair
// per block into partials, a fixed tree of k_sum_pairs reduces them into
// result[0..1].  partials holds 2 * partial_pairs_capacity(2^nq) doubles;
// every call writes what it reads.  With timed != nullptr the launches are
// timed on that context.
hipError_t launch_norm_tree(const double* re, const double* im, uint32_t nq, const Sel& sel, uint64_t split_bit,
                            double* partials, double* result, hipStream_t st, suqa_hip_ctx* timed, const char* tag,
                            double alg_bytes);
// A gate whose qubits lie in the active gc_mask, with the reference's dense
// loop bodies on coinciding slots (kernels.hip, k_overlap): selection
// (i0 & (gc | gate bits)) == gc, slots i0 | {0, m1, m2, m1 | m2}.
constexpr int kOvH = 0, kOvSwap = 1, kOvPauli1 = 2, kOvPauli2 = 3;
hipError_t launch_overlap(suqa_hip_ctx* c, int op, const Sel& sel, uint64_t m1, uint64_t m2, uint32_t p1, uint32_t p2,
                          double cs, double sn);
hipError_t launch_collapse(suqa_hip_ctx* c, uint64_t qbit, uint32_t keep, double scale);
// d_result[0..2] = sum re, sum im, sum i |psi_i|^2 over the whole state.
hipError_t launch_checksum(suqa_hip_ctx* c);
// Measurements of a state of at most 2^kMeasSmallBits amplitudes in one
// workgroup: for k < n in order, p0 and p1 of qubit qs[k], the outcome
// c = (rds[k] < p1) into ctx->d_flags[k], the collapse scaled by 1/sqrt(p_c),
// and when bit k of `resets` is set and c = 1, X on qs[k] over the indices
// with the gc bits set; with xqs, then X on qubit xqs[k] over the same
// indices where the device flag xflags[k] reads 1.  n <= kMeasMax.
constexpr uint32_t kMeasSmallBits = 14, kMeasMax = 16;
hipError_t launch_measure_small(suqa_hip_ctx* c, const uint32_t* qs, const double* rds, uint32_t n, uint32_t resets,
                                uint32_t gc, uint32_t* out = nullptr,   // out: device outcomes (d_flags when null)
                                const uint32_t* xqs = nullptr, const uint32_t* xflags = nullptr);
// X on qubit qs[k] over the indices with the gc bits set where the device
// outcome flags[k] is 1, k < n <= kMeasMax, in order (one launch).
hipError_t launch_x_if(suqa_hip_ctx* c, const uint32_t* qs, const uint32_t* flags, uint32_t n, uint32_t gc);
// W of qms.cuh:220-242 over the acceptance pairs (i, i without the top bit):
// coefficients (fs1, fs2) at d_tab[2 (Eo L + En)], En = (i & wmask) >> b0 and
// Eo = (i & omask) >> ob0 (omask = 0: Eo = 0, d_tab is the row of a classical
// E_old).
hipError_t launch_qms_W(suqa_hip_ctx* c, const double* d_tab, uint32_t wmask, uint32_t b0, uint32_t omask,
                        uint32_t ob0, uint32_t L);
hipError_t launch_copy(suqa_hip_ctx* c, void* dst, const void* src, uint64_t nbytes);
// Number of (sum, sum) partial pairs launch_norm needs for a 2^nq state.
uint64_t partial_pairs_capacity(uint64_t N);

// One-gate sweeps (kernels.hip), called directly or by the pass scheduler.
hipError_t launch_pair_now(suqa_hip_ctx* c, int op, const Sel& sel, uint64_t xmask, const PauliCoef& pc,
                           const char* tag, double alg_bytes);
hipError_t launch_phase_now(suqa_hip_ctx* c, const Sel& sel, double cs, double sn, uint64_t zmask,
                            const char* tag, double alg_bytes);
// Applies the queued gates in tile passes (tiled.hip); a no-op when empty.
hipError_t flush(suqa_hip_ctx* c);
// The flush a queue of auto_flush_at(c) operations triggers: the passes of the
// queue but its last ones, whose gates stay queued to be packed with the next
// operations (tiled.hip).
hipError_t flush_auto(suqa_hip_ctx* c);
// States of at most 2^kTinyProgBits amplitudes that are one pass tile
// (tile_bits >= nq) are latency-bound: measurements, resets and apply_W queue
// beside the gates, and a flush runs the queue as programs of up to kProgMax
// operations, one single-workgroup launch each (k_tiny_prog, tiled.hip), with
// the one-gate sweeps' arithmetic and launch_measure_small's reduction order
// (bit-identical results).  Outcomes go to a ring of kSlots device slots.
constexpr uint32_t kTinyProgBits = 10, kProgMax = 96, kSlots = 1024;
inline bool tiny_prog(const suqa_hip_ctx* c) {
    return !c->f32 && c->tile_bits >= c->nq && c->nq >= 1 && c->nq <= kTinyProgBits && c->d_slots;
}
// n consecutive outcome slots (n <= kSlots); the ring wraps to 0 when fewer
// are left before its end.  A slot keeps its outcome until kSlots later
// slots have been handed out.
inline uint32_t alloc_slots(suqa_hip_ctx* c, uint32_t n) {
    if (c->slot_next + n > kSlots) c->slot_next = 0;
    const uint32_t b = c->slot_next;
    c->slot_next += n;
    return b;
}
// Queues the measurements of qs[0..n) (draws rds, consecutive slots from
// `slot`), resets when `reset` (X over the gc bits when the outcome is 1).
void queue_measure(suqa_hip_ctx* c, const uint32_t* qs, const double* rds, uint32_t n, bool reset, uint32_t gc,
                   uint32_t slot);
constexpr uint32_t kMinTileBits = 1, kMaxTileBits = 12;
constexpr uint32_t kMaxTileBits32 = 13;   // float tiles (64 KiB of LDS)
constexpr uint32_t kPassLowFixed = 4;    // qubits 0..3 are in every pass tile (one 128-byte line)
constexpr uint32_t kPassLowFixed32 = 5;  // the same for float states: qubits 0..4
constexpr uint32_t kPassMaxGates = 32;   // gates per pass
// The selection a pass must honour for gate d.  A Pauli string's sweep
// enumerates each pair (i, i ^ X) once, from the element whose lowest flipped
// qubit is 0 (that qubit in M at value 0); its arithmetic treats both elements
// of a pair alike (pair_out_par), so when no other flipped qubit is selected
// the pass may pick each pair's first element by a tile coordinate instead,
// and the qubit needs no coordinate of its own.
inline uint64_t pass_sel(const GateDesc& d) {
    if (d.kind != OP_PAULI) return d.M;
    const uint64_t p = d.X & (0ull - d.X);
    if ((d.M & d.X) == p && (d.V & p) == 0ull) return d.M & ~p;
    return d.M;
}

// The tile coordinates that a pass of gates needs above its low qubits.  Span
// tiles: the GF(2) span of the gates' flip masks (the part above the low
// qubits), where every selection qubit inside the span's support gets a
// coordinate of its own (a unit vector: a selection must read one tile
// coordinate), so dims = |sel & supp| + rank{flip & ~(sel & supp)}.  Qubit
// tiles (span off): one coordinate per flipped qubit, dims = |supp|.  For
// gates that flip single qubits both counts are the same.
struct PassSpace {
    uint64_t low = 0;          // the low qubits (always whole tile coordinates)
    bool span = true;
    uint64_t sel = 0, supp = 0;
    uint32_t nflips = 0;
    uint64_t flips[kPassMaxGates] = {};
    static uint32_t rank_of(const uint64_t* v, uint32_t n, uint64_t clear) {
        uint64_t basis[64] = {};
        uint32_t r = 0;
        for (uint32_t i = 0; i < n; ++i) {
            uint64_t x = v[i] & ~clear;
            while (x) {
                const uint32_t hb = 63u - (uint32_t)__builtin_clzll(x);
                if (!basis[hb]) {
                    basis[hb] = x;
                    ++r;
                    break;
                }
                x ^= basis[hb];
            }
        }
        return r;
    }
    uint32_t dims() const {
        if (!span) return (uint32_t)__builtin_popcountll(supp);
        const uint64_t q = sel & supp;
        return (uint32_t)__builtin_popcountll(q) + rank_of(flips, nflips, q);
    }
    // Adds a gate with flip mask X and selection mask M if the pass then needs
    // at most `room` coordinates above the low qubits (and holds fewer than
    // kPassMaxGates flipping gates); false leaves the space unchanged.
    bool add(uint64_t X, uint64_t M, uint32_t room) {
        const uint64_t h = X & ~low;
        if (h && nflips >= kPassMaxGates) return false;
        const uint64_t sel0 = sel, supp0 = supp;
        sel |= M & ~low;
        supp |= h;
        if (h) flips[nflips++] = h;
        if (dims() <= room) return true;
        sel = sel0;
        supp = supp0;
        if (h) --nflips;
        return false;
    }
};

// SUQA_HIP_SPAN_TILES=0 in the environment turns span tiles off (capi.cpp)
bool span_tiles_env();
// run_paulis' order of an op list with gate passes on (capi.cpp)
std::vector<uint32_t> pauli_pass_order(const suqa_hip_ctx* c, const suqa_hip_pauli_op* ops, uint32_t n);
constexpr size_t kAutoFlush = 256;   // queued gates that trigger a flush (keeps the GPU busy)
constexpr size_t kAutoFlush32 = 1024;   // the same for queues that pack commuting gates (a longer reorder span)
// Queue length that triggers a flush: a full program for small states (the GPU
// runs it while the host queues the next operations), else the above.
inline size_t auto_flush_at(const suqa_hip_ctx* c) {
    if (tiny_prog(c)) return kProgMax;
    return (c->f32 || c->gate_reorder) ? kAutoFlush32 : kAutoFlush;
}

// the message suqa_hip_last_error() returns on this thread (capi.cpp)
void set_last_error(const std::string& msg);

// A gate queue over an FP32 state of 2^nq amplitudes (re, im on `stream`),
// for the FP32 engine's gate passes: the FP64 gate entry points
// (suqa_hip_run_gates, suqa_hip_run_paulis, suqa_hip_apply_qft_inverse) accept
// it and queue their gates, flush() applies them as float tile passes of
// tile_bits qubits.  Nothing is allocated on the device.  (capi.cpp)
suqa_hip_ctx* f32_gate_queue(int device, hipStream_t stream, uint32_t nq, float* re, float* im);
void f32_gate_queue_free(suqa_hip_ctx* q);
void timing_flush(suqa_hip_ctx* c);

// timing hooks (capi.cpp)
void timing_begin(suqa_hip_ctx* c, const char* tag, double alg_bytes);
void timing_end(suqa_hip_ctx* c);

}  // namespace sqe
