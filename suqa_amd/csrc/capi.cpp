// capi.cpp -- the extern "C" entry points of include/suqa_hip.h.
//
// Host side of each reference API call: build the reference's bit masks
// (suqa.cu), evaluate sincos/exp/sqrt with the host libm exactly where the
// reference does, turn the predicate into a dense selection and launch one
// sweep.  No CPU fallback: every amplitude update happens in a HIP kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "engine.hpp"

using sqe::PauliCoef;
using sqe::Sel;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define SQE_HIP(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail(SUQA_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

#define SQE_CTX(c)                                                                             \
    do {                                                                                       \
        if (!(c) || (!(c)->d_data && !(c)->f32)) return fail(SUQA_ERR_ARG, "null or cleared context"); \
    } while (0)

uint32_t popc32(uint32_t x) { return (uint32_t)__builtin_popcount(x); }

int read_result(suqa_hip_ctx* c, int n) {
    if (c->capturing) return fail(SUQA_ERR_ARG, "a value cannot be read back while a graph is being captured");
    SQE_HIP(hipMemcpyAsync(c->h_result, c->d_result, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    SQE_HIP(hipStreamSynchronize(c->stream));
    return SUQA_OK;
}

// Tile bits of a new context: SUQA_HIP_TILE_BITS (0 = one sweep per gate,
// 1..12) when set to a valid value, else kDefaultTileBits (the fastest size
// measured on the 30-qubit random circuit).
constexpr uint32_t kDefaultTileBits = 11;
uint32_t default_tile_bits() {
    const char* e = std::getenv("SUQA_HIP_TILE_BITS");
    if (!e || !*e) return kDefaultTileBits;
    char* end = nullptr;
    const long v = std::strtol(e, &end, 10);
    if (*end || v < 0 || v > (long)sqe::kMaxTileBits) return kDefaultTileBits;
    return (uint32_t)v;
}

void drop_graphs(suqa_hip_ctx* c) {
    for (auto& g : c->graphs)
        if (g) (void)hipGraphExecDestroy(g);
    c->graphs.clear();
}

int free_state(suqa_hip_ctx* c) {
    c->queue.clear();
    drop_graphs(c);   // they hold the state's addresses
    if (c->d_data) SQE_HIP(hipFree(c->d_data));
    if (c->d_partials) SQE_HIP(hipFree(c->d_partials));
    c->d_partials = nullptr;
    c->d_data = c->d_re = c->d_im = nullptr;
    c->nq = 0;
    c->N = 0;
    return SUQA_OK;
}

// exp(i phase P) with P given by X/Y/Z masks (qubits already validated).
int pauli_masks(suqa_hip_ctx* c, uint32_t xm, uint32_t ym, uint32_t zm, double phase, uint32_t gc) {
    double s, cs;
    sincos(phase, &s, &cs);
    const uint64_t xy = (uint64_t)(xm | ym), yz = (uint64_t)(ym | zm);
    const uint32_t ny = popc32(ym);
    Sel sel;
    if (xy == 0) {
        // diagonal: psi' = (c + i (-1)^{par(i&Z)} s) psi  (util_rotate2 form)
        if (!sqe::make_sel(c->nq, gc, gc, &sel)) return SUQA_OK;
        SQE_HIP(sqe::launch_phase(c, sel, cs, s, yz, "pauli_diag", 32.0 * (double)sel.count));
        return SUQA_OK;
    }
    const uint64_t pivot = xy & (~xy + 1ull);
    if (!sqe::make_sel(c->nq, (uint64_t)gc | pivot, gc, &sel)) return SUQA_OK;
    PauliCoef pc;
    pc.c = cs;
    pc.s = s;
    pc.yz = yz;
    pc.ny_odd = ny & 1u;
    pc.base_imag = (ny + 1u) & 1u;
    pc.base_neg = ((ny + 1u) >> 1) & 1u;
    SQE_HIP(sqe::launch_pair(c, sqe::OP_PAULI, sel, xy, pc, "pauli_pair", 64.0 * (double)sel.count));
    return SUQA_OK;
}

int check_qubits(const suqa_hip_ctx* c, const uint32_t* qs, uint32_t n, uint32_t gc, bool gc_overlap_ok = false) {
    uint32_t seen = 0;
    for (uint32_t k = 0; k < n; ++k) {
        if (qs[k] >= c->nq) return fail(SUQA_ERR_ARG, "qubit index out of range");
        const uint32_t b = 1u << qs[k];
        if (seen & b) return fail(SUQA_ERR_ARG, "repeated qubit in a Pauli string");
        if ((gc & b) && !gc_overlap_ok) return fail(SUQA_ERR_ARG, "Pauli rotation on a qubit of the active gc_mask");
        seen |= b;
    }
    return SUQA_OK;
}

// crm of the QFT (suqa.cu:620-633).  The reference fills its phase table once,
// sized by the first register; entry m is 2 pi / 2^m whatever the size, so a
// larger register later extends the table instead of reading past its end.
int qft_crm(suqa_hip_ctx* c, uint32_t qc, uint32_t qt, int m, int nqubits, uint32_t gc) {
    const size_t idx = (size_t)(m > 0 ? m : -m);
    if (!c->rphase_filled || idx >= c->rphase.size()) {
        const size_t sz = std::max((size_t)nqubits, idx) + 1;
        c->rphase.assign(sz, 0.0);
        uint32_t cc = 1;
        for (size_t i = 0; i < sz; ++i) {
            c->rphase[i] = (2.0 * M_PI / (double)cc);
            cc <<= 1;
        }
        c->rphase_filled = true;
    }
    double rphase = c->rphase[idx];
    if (m <= 0) rphase *= -1;
    const uint32_t one = 1u;
    return suqa_hip_apply_mcu1(c, &qc, &one, 1, qt, rphase, gc);
}

}  // namespace

// ------------------------------------------------------------------ timing
namespace sqe {

void set_last_error(const std::string& msg) { g_err = msg; }

void timing_flush(suqa_hip_ctx* c) {
    if (c->recs.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (auto& r : c->recs) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.e0, r.e1) == hipSuccess) c->stats[r.stat].total_ms += ms;
        c->evpool.push_back(r.e0);
        c->evpool.push_back(r.e1);
    }
    c->recs.clear();
}

static hipEvent_t get_event(suqa_hip_ctx* c) {
    if (!c->evpool.empty()) {
        hipEvent_t e = c->evpool.back();
        c->evpool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void timing_begin(suqa_hip_ctx* c, const char* tag, double alg_bytes) {
    if (!c->timing || c->capturing) return;
    if (c->recs.size() >= 4096) timing_flush(c);
    int idx = -1;
    for (size_t k = 0; k < c->stats.size(); ++k)
        if (c->stats[k].name == tag) idx = (int)k;
    if (idx < 0) {
        c->stats.push_back(KStat{tag, 0, 0.0, 0.0});
        idx = (int)c->stats.size() - 1;
    }
    c->stats[idx].launches++;
    c->stats[idx].alg_bytes += alg_bytes;
    suqa_hip_ctx::Rec r{idx, get_event(c), get_event(c)};
    (void)hipEventRecord(r.e0, c->stream);
    c->recs.push_back(r);
}

void timing_end(suqa_hip_ctx* c) {
    if (!c->timing || c->capturing || c->recs.empty()) return;
    (void)hipEventRecord(c->recs.back().e1, c->stream);
}

// Span tiles for new contexts and gate queues unless SUQA_HIP_SPAN_TILES=0.
bool span_tiles_env() {
    const char* e = std::getenv("SUQA_HIP_SPAN_TILES");
    return !(e && std::strcmp(e, "0") == 0);
}

suqa_hip_ctx* f32_gate_queue(int device, hipStream_t stream, uint32_t nq, float* re, float* im) {
    if (nq < 1 || nq > 33 || !re || !im) return nullptr;
    auto* q = new suqa_hip_ctx();
    q->device = device;
    q->stream = stream;   // the FP32 engine's stream (not owned)
    q->nq = nq;
    q->N = 1ull << nq;
    q->f32 = true;
    q->f32_re = re;
    q->f32_im = im;
    q->tile_bits = kMaxTileBits32 - 1;
    const char* e = std::getenv("SUQA_HIP_PAULI_ORDER");
    q->pauli_reorder = !(e && std::strcmp(e, "0") == 0);
    const char* e2 = std::getenv("SUQA_HIP_GATE_ORDER");
    q->gate_reorder = !(e2 && std::strcmp(e2, "0") == 0);
    q->span_tiles = span_tiles_env();
    return q;
}

void f32_gate_queue_free(suqa_hip_ctx* q) {
    if (!q) return;
    timing_flush(q);
    for (auto e : q->evpool) (void)hipEventDestroy(e);
    delete q;
}

}  // namespace sqe

namespace {

// Symplectic form of a controlled Pauli rotation: flipped qubits (X|Y),
// phase qubits (Y|Z), controls.
struct PauliSym {
    uint32_t x, z, g;
};

PauliSym pauli_sym(const suqa_hip_pauli_op& o) {
    PauliSym p{0u, 0u, o.gc_mask};
    for (uint32_t k = 0; k < o.w && k < 8; ++k) {
        const uint32_t b = o.qs[k] < 32 ? 1u << o.qs[k] : 0u;
        if (o.types[k] == SUQA_PAULI_X || o.types[k] == SUQA_PAULI_Y) p.x |= b;
        if (o.types[k] == SUQA_PAULI_Y || o.types[k] == SUQA_PAULI_Z) p.z |= b;
    }
    return p;
}

// Two controlled rotations exp(i a P) and exp(i b Q) commute when neither
// flips a control of the other (their projectors then commute with the other
// rotation) and P, Q commute (an even number of positions where they differ
// and neither is I).
bool commute(const PauliSym& a, const PauliSym& b) {
    if ((a.g & b.x) || (b.g & a.x)) return false;
    return (__builtin_popcount((a.x & b.z) ^ (a.z & b.x)) & 1) == 0;
}

// The order in which run_paulis queues a list with gate passes on: passes
// packed greedily, an op moving ahead only past ops it commutes with, so the
// product is the list's.  A pass is a run of at most `max_pass` ops whose
// flip masks fit, with qubits 0..low-1, in a tile of `t` coordinates
// (PassSpace: the span of the flip masks, or the flipped qubits), as the
// flush forms it (tiled.hip), so the flush cuts the reordered queue into
// exactly these passes.  Lookahead over `window` ops.
std::vector<uint32_t> pass_order(const suqa_hip_pauli_op* ops, uint32_t n, uint32_t t, uint32_t low, bool span,
                                 uint32_t max_pass, uint32_t window) {
    std::vector<PauliSym> sym(n);
    for (uint32_t k = 0; k < n; ++k) sym[k] = pauli_sym(ops[k]);
    std::vector<uint32_t> rest(n), order;
    for (uint32_t k = 0; k < n; ++k) rest[k] = k;
    order.reserve(n);
    const uint32_t f = std::min(low, t);
    while (!rest.empty()) {
        sqe::PassSpace ps;
        ps.low = (1ull << f) - 1ull;
        ps.span = span;
        uint32_t count = 0;
        bool took = true;
        while (took && count < max_pass && !rest.empty()) {
            took = false;
            const size_t lim = std::min<size_t>(rest.size(), window);
            for (size_t j = 0; j < lim; ++j) {
                const PauliSym& s = sym[rest[j]];
                bool ok = true;   // commutes with every op it would move ahead of
                for (size_t i = 0; i < j && ok; ++i) ok = commute(sym[rest[i]], s);
                if (!ok || !ps.add(s.x, s.g, t - f)) continue;
                ++count;
                order.push_back(rest[j]);
                rest.erase(rest.begin() + (ptrdiff_t)j);
                took = true;
                break;
            }
        }
        if (count == 0) {   // the first op alone (its flips exceed the tile)
            order.push_back(rest.front());
            rest.erase(rest.begin());
        }
    }
    return order;
}

}  // namespace

namespace sqe {
std::vector<uint32_t> pauli_pass_order(const suqa_hip_ctx* c, const suqa_hip_pauli_op* ops, uint32_t n) {
    return pass_order(ops, n, std::min(c->tile_bits, c->nq), c->f32 ? kPassLowFixed32 : kPassLowFixed, c->span_tiles,
                      kPassMaxGates, 128);
}
}  // namespace sqe

// ------------------------------------------------------------ entry points
extern "C" {

const char* suqa_hip_last_error(void) { return g_err.c_str(); }

int suqa_hip_allocate_state(suqa_hip_ctx* c, uint32_t nq) {
    if (!c) return fail(SUQA_ERR_ARG, "null context");
    if (nq < 1 || nq > 32) return fail(SUQA_ERR_ARG, "nq must be in [1, 32]");
    if (c->d_data && c->nq == nq) return suqa_hip_init_state(c);
    int rc = free_state(c);
    if (rc) return rc;
    const uint64_t N = 1ull << nq;
    void* p = nullptr;
    if (hipMalloc(&p, 2 * N * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        return fail(SUQA_ERR_NOMEM, "hipMalloc of the state failed");
    }
    c->d_data = static_cast<double*>(p);
    c->d_re = c->d_data;
    c->d_im = c->d_data + N;
    c->nq = nq;
    c->N = N;
    if (hipMalloc(&p, sizeof(double) * 2 * sqe::partial_pairs_capacity(N)) != hipSuccess) {
        (void)hipGetLastError();
        free_state(c);
        return fail(SUQA_ERR_NOMEM, "hipMalloc of the reduction partials failed");
    }
    c->d_partials = static_cast<double*>(p);
    return suqa_hip_init_state(c);
}

int suqa_hip_setup(uint32_t nq, suqa_hip_ctx** out) {
    if (!out) return fail(SUQA_ERR_ARG, "null output pointer");
    *out = nullptr;
    if (nq < 1 || nq > 32) return fail(SUQA_ERR_ARG, "nq must be in [1, 32]");
    auto* c = new suqa_hip_ctx();
    // every failure below frees what was made so far (suqa_hip_clear) and
    // keeps the failing call's message
#define SETUP_HIP(call)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess) {                                                                \
            const std::string msg_ = std::string(#call) + ": " + hipGetErrorString(e_);        \
            suqa_hip_clear(c);                                                                 \
            return fail(SUQA_ERR_HIP, msg_);                                                   \
        }                                                                                      \
    } while (0)
    SETUP_HIP(hipGetDevice(&c->device));
    int ncu = 0;
    SETUP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
    c->num_cu = ncu > 0 ? ncu : 256;
    SETUP_HIP(hipStreamCreate(&c->stream));
    c->tile_bits = default_tile_bits();
    {
        const char* e = std::getenv("SUQA_HIP_PAULI_ORDER");
        c->pauli_reorder = !(e && std::strcmp(e, "0") == 0);
    }
    c->span_tiles = sqe::span_tiles_env();
    void* p = nullptr;
    SETUP_HIP(hipMalloc(&p, sizeof(double) * 4));
    c->d_result = static_cast<double*>(p);
    SETUP_HIP(hipHostMalloc(&p, sizeof(double) * 4, hipHostMallocDefault));
    c->h_result = static_cast<double*>(p);
    SETUP_HIP(hipMalloc(&p, sizeof(uint32_t) * sqe::kMeasMax));
    c->d_flags = static_cast<uint32_t*>(p);
    SETUP_HIP(hipHostMalloc(&p, sizeof(uint32_t) * sqe::kMeasMax, hipHostMallocDefault));
    c->h_flags = static_cast<uint32_t*>(p);
    SETUP_HIP(hipMalloc(&p, sizeof(uint32_t) * sqe::kSlots));
    c->d_slots = static_cast<uint32_t*>(p);
    SETUP_HIP(hipMemsetAsync(c->d_slots, 0, sizeof(uint32_t) * sqe::kSlots, c->stream));
    SETUP_HIP(hipHostMalloc(&p, sizeof(uint32_t) * sqe::kSlots, hipHostMallocDefault));
    c->h_slots = static_cast<uint32_t*>(p);
#undef SETUP_HIP
    int rc = suqa_hip_allocate_state(c, nq);
    if (rc) {
        std::string keep = g_err;
        suqa_hip_clear(c);
        return fail(rc, keep);
    }
    *out = c;
    return SUQA_OK;
}

int suqa_hip_clear(suqa_hip_ctx* c) {
    if (!c) return SUQA_OK;
    if (c->capturing) {
        hipGraph_t g = nullptr;
        if (hipStreamEndCapture(c->stream, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
        c->capturing = false;
    }
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& r : c->recs) {
        (void)hipEventDestroy(r.e0);
        (void)hipEventDestroy(r.e1);
    }
    for (auto e : c->evpool) (void)hipEventDestroy(e);
    free_state(c);
    if (c->d_result) (void)hipFree(c->d_result);
    if (c->h_result) (void)hipHostFree(c->h_result);
    if (c->d_flags) (void)hipFree(c->d_flags);
    if (c->h_flags) (void)hipHostFree(c->h_flags);
    if (c->d_slots) (void)hipFree(c->d_slots);
    if (c->h_slots) (void)hipHostFree(c->h_slots);
    if (c->d_wtab) (void)hipFree(c->d_wtab);
    if (c->d_wrow) (void)hipFree(c->d_wrow);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SUQA_OK;
}

int suqa_hip_init_state(suqa_hip_ctx* c) {
    SQE_CTX(c);
    SQE_HIP(sqe::launch_init(c));
    return SUQA_OK;
}

int suqa_hip_vnorm(suqa_hip_ctx* c, double* norm) {
    SQE_CTX(c);
    Sel sel;
    sqe::make_sel(c->nq, 0, 0, &sel);
    SQE_HIP(sqe::launch_norm(c, sel, 0, "norm", 16.0 * (double)c->N));
    int rc = read_result(c, 1);
    if (rc) return rc;
    if (norm) *norm = std::sqrt(c->h_result[0]);
    return SUQA_OK;
}

int suqa_hip_checksum(suqa_hip_ctx* c, double* out3) {
    SQE_CTX(c);
    if (!out3) return fail(SUQA_ERR_ARG, "null argument");
    SQE_HIP(sqe::launch_checksum(c));
    int rc = read_result(c, 3);
    if (rc) return rc;
    out3[0] = c->h_result[0];
    out3[1] = c->h_result[1];
    out3[2] = c->h_result[2] / (double)c->N;
    return SUQA_OK;
}

int suqa_hip_vnormalize(suqa_hip_ctx* c) {
    double v = 0.0;
    int rc = suqa_hip_vnorm(c, &v);
    if (rc) return rc;
    SQE_HIP(sqe::launch_scale(c, 1. / v));
    return SUQA_OK;
}

// X/Y on q: pairs (j, j|T) with the reference predicate (i & (gc|T)) == gc|T on i = j|T.
static int pair_1q(suqa_hip_ctx* c, int op, uint32_t q, uint32_t gc, const char* tag) {
    SQE_CTX(c);
    if (q >= c->nq) return fail(SUQA_ERR_ARG, "qubit index out of range");
    const uint64_t T = 1ull << q;
    const uint64_t M = (uint64_t)gc | T;
    Sel sel;
    if (!sqe::make_sel(c->nq, M, M & ~T, &sel)) return SUQA_OK;
    PauliCoef pc{};
    SQE_HIP(sqe::launch_pair(c, op, sel, T, pc, tag, 64.0 * (double)sel.count));
    return SUQA_OK;
}

int suqa_hip_apply_x(suqa_hip_ctx* c, uint32_t q, uint32_t gc) { return pair_1q(c, sqe::OP_SWAP, q, gc, "swap_pair"); }
int suqa_hip_apply_y(suqa_hip_ctx* c, uint32_t q, uint32_t gc) { return pair_1q(c, sqe::OP_Y, q, gc, "y_pair"); }

int suqa_hip_apply_h(suqa_hip_ctx* c, uint32_t q, uint32_t gc) {
    SQE_CTX(c);
    if (q >= c->nq) return fail(SUQA_ERR_ARG, "qubit index out of range");
    const uint64_t T = 1ull << q;
    Sel sel;
    if (gc & (1u << q)) {   // the reference's loop on coinciding slots (suqa_cpu.hpp:186-198)
        if (sqe::make_sel(c->nq, (uint64_t)gc | T, gc, &sel))
            SQE_HIP(sqe::launch_overlap(c, sqe::kOvH, sel, T, 0, 0, 0, 0.0, 0.0));
        return SUQA_OK;
    }
    if (!sqe::make_sel(c->nq, (uint64_t)gc | T, gc, &sel)) return SUQA_OK;
    PauliCoef pc{};
    SQE_HIP(sqe::launch_pair(c, sqe::OP_H, sel, T, pc, "h_pair", 64.0 * (double)sel.count));
    return SUQA_OK;
}

int suqa_hip_apply_u1(suqa_hip_ctx* c, uint32_t q, uint32_t q_mask, double phase, uint32_t gc) {
    SQE_CTX(c);
    if (q >= c->nq) return fail(SUQA_ERR_ARG, "qubit index out of range");
    const uint32_t qmask = gc | (q_mask << q);       // suqa.cu:272
    const uint32_t glob = gc | (1u << q);            // suqa_cpu.hpp:211
    double s, cs;
    sincos(phase, &s, &cs);
    Sel sel;
    if (!sqe::make_sel(c->nq, glob, qmask, &sel)) return SUQA_OK;
    SQE_HIP(sqe::launch_phase(c, sel, cs, s, 0, "phase", 32.0 * (double)sel.count));
    return SUQA_OK;
}

int suqa_hip_apply_mcx(suqa_hip_ctx* c, const uint32_t* ctrls, const uint32_t* vals, uint32_t nctrl,
                       uint32_t target, uint32_t gc) {
    SQE_CTX(c);
    if (target >= c->nq) return fail(SUQA_ERR_ARG, "target index out of range");
    if (nctrl && !ctrls) return fail(SUQA_ERR_ARG, "null control list");
    uint32_t mask = gc, mqs = gc;                    // suqa.cu:290-292, 308-310, 326-332
    for (uint32_t k = 0; k < nctrl; ++k) {
        if (ctrls[k] >= c->nq) return fail(SUQA_ERR_ARG, "control index out of range");
        mask |= 1u << ctrls[k];
        if (!vals || vals[k]) mqs |= 1u << ctrls[k];
    }
    const uint64_t T = 1ull << target;               // suqa_cpu.hpp:239-240
    const uint64_t M = (uint64_t)mask | T, Vup = (uint64_t)mqs | T;
    Sel sel;
    if (!sqe::make_sel(c->nq, M, Vup & ~T, &sel)) return SUQA_OK;
    PauliCoef pc{};
    SQE_HIP(sqe::launch_pair(c, sqe::OP_SWAP, sel, T, pc, "swap_pair", 64.0 * (double)sel.count));
    return SUQA_OK;
}

int suqa_hip_apply_mcu1(suqa_hip_ctx* c, const uint32_t* ctrls, const uint32_t* vals, uint32_t nctrl,
                        uint32_t target, double phase, uint32_t gc) {
    SQE_CTX(c);
    if (target >= c->nq) return fail(SUQA_ERR_ARG, "target index out of range");
    if (nctrl && !ctrls) return fail(SUQA_ERR_ARG, "null control list");
    uint32_t mask = (1u << target) | gc, mqs = (1u << target) | gc;   // suqa.cu:347-349, 367-373
    for (uint32_t k = 0; k < nctrl; ++k) {
        if (ctrls[k] >= c->nq) return fail(SUQA_ERR_ARG, "control index out of range");
        mask |= 1u << ctrls[k];
        if (!vals || vals[k]) mqs |= 1u << ctrls[k];
    }
    double s, cs;
    sincos(phase, &s, &cs);
    Sel sel;
    if (!sqe::make_sel(c->nq, mask, mqs, &sel)) return SUQA_OK;
    SQE_HIP(sqe::launch_phase(c, sel, cs, s, 0, "phase", 32.0 * (double)sel.count));
    return SUQA_OK;
}

int suqa_hip_apply_swap(suqa_hip_ctx* c, uint32_t q1, uint32_t q2, uint32_t gc) {
    SQE_CTX(c);
    if (q1 >= c->nq || q2 >= c->nq) return fail(SUQA_ERR_ARG, "qubit index out of range");
    if (q1 == q2) return SUQA_OK;                    // the reference swaps each amplitude with itself
    const uint32_t m1 = 1u << q1, m2 = 1u << q2;
    // suqa_cpu.hpp:303-309: (i & (gc|m1|m2)) == gc, swap i|m1 <-> i|m2
    Sel sel;
    if (gc & (m1 | m2)) {   // on coinciding slots
        if (sqe::make_sel(c->nq, (uint64_t)(gc | m1 | m2), gc, &sel))
            SQE_HIP(sqe::launch_overlap(c, sqe::kOvSwap, sel, m1, m2, 0, 0, 0.0, 0.0));
        return SUQA_OK;
    }
    if (!sqe::make_sel(c->nq, (uint64_t)(gc | m1 | m2), (uint64_t)(gc | m1), &sel)) return SUQA_OK;
    PauliCoef pc{};
    SQE_HIP(sqe::launch_pair(c, sqe::OP_SWAP, sel, (uint64_t)(m1 | m2), pc, "swap_pair",
                             64.0 * (double)sel.count));
    return SUQA_OK;
}

int suqa_hip_apply_phase_list(suqa_hip_ctx* c, uint32_t q0, uint32_t q_size, const double* phases,
                              uint32_t nphases, uint32_t gc) {
    SQE_CTX(c);
    if (!(q_size > 0u && q_size <= 7u && nphases == (1u << q_size)) || !phases)
        return fail(SUQA_ERR_ARG, "in suqa::apply_phase_list(): invalid q_size or phases.size()");
    if (q0 >= c->nq) return fail(SUQA_ERR_ARG, "qubit index out of range");
    double cs[128], sn[128];
    for (uint32_t k = 0; k < nphases; ++k) sincos(phases[k], &sn[k], &cs[k]);
    Sel sel;
    if (!sqe::make_sel(c->nq, gc, gc, &sel)) return SUQA_OK;
    SQE_HIP(sqe::launch_phase_list(c, sel, q0, (1u << q_size) - 1u, cs, sn));
    return SUQA_OK;
}

int suqa_hip_apply_pauli_TP_rotation(suqa_hip_ctx* c, const uint32_t* qs, const uint32_t* types, uint32_t n,
                                     double phase, uint32_t gc) {
    SQE_CTX(c);
    if (n == 0 || n > 3)
        return fail(SUQA_ERR_UNSUPPORTED, "ERROR: unimplemented pauli TP rotation in the selected configuration");
    if (!qs || !types) return fail(SUQA_ERR_ARG, "null argument");
    int rc = check_qubits(c, qs, n, gc, n <= 2);
    if (rc) return rc;
    uint32_t t[3], q[3];
    for (uint32_t k = 0; k < n; ++k) {
        t[k] = types[k];
        q[k] = qs[k];
    }
    auto sw = [&](int a, int b) {
        std::swap(t[a], t[b]);
        std::swap(q[a], q[b]);
    };
    uint32_t xm = 0, ym = 0, zm = 0;
    auto add = [&](uint32_t type, uint32_t qq) {
        if (type == SUQA_PAULI_X) xm |= 1u << qq;
        else if (type == SUQA_PAULI_Y) ym |= 1u << qq;
        else if (type == SUQA_PAULI_Z) zm |= 1u << qq;
    };
    uint32_t over = 0;   // gate qubits inside the active gc_mask
    for (uint32_t k = 0; k < n; ++k) over |= gc & (1u << q[k]);
    if (n == 1) {
        if (t[0] < 1 || t[0] > 3) return SUQA_OK;    // suqa_cpu.hpp:360-370: no matching case
        if (over) {   // the reference's loop on coinciding slots
            double sph, cph;
            sincos(phase, &sph, &cph);
            const uint64_t m1 = 1ull << q[0];
            Sel sel;
            if (sqe::make_sel(c->nq, (uint64_t)gc | m1, gc, &sel))
                SQE_HIP(sqe::launch_overlap(c, sqe::kOvPauli1, sel, m1, 0, t[0], 0, cph, sph));
            return SUQA_OK;
        }
        add(t[0], q[0]);
    } else if (n == 2) {
        if (t[0] > t[1]) sw(0, 1);                   // suqa.cu:506-509
        const bool known = t[0] >= 1 && t[1] <= 3 && t[0] <= t[1];
        if (!known) return SUQA_OK;                  // suqa_cpu.hpp:422-440: no matching case
        if (over) {   // the reference's loop on coinciding slots
            double sph, cph;
            sincos(phase, &sph, &cph);
            const uint64_t m1 = 1ull << q[0], m2 = 1ull << q[1];
            Sel sel;
            if (sqe::make_sel(c->nq, (uint64_t)gc | m1 | m2, gc, &sel))
                SQE_HIP(sqe::launch_overlap(c, sqe::kOvPauli2, sel, m1, m2, t[0], t[1], cph, sph));
            return SUQA_OK;
        }
        if (t[0] == SUQA_PAULI_X && t[1] == SUQA_PAULI_Y) {
            // suqa_cpu.hpp:431-433: the reference's X(q1)Y(q2) block rotates
            // about Y on q1 and X on q2.
            add(SUQA_PAULI_Y, q[0]);
            add(SUQA_PAULI_X, q[1]);
        } else {
            add(t[0], q[0]);
            add(t[1], q[1]);
        }
    } else {
        if (t[0] > t[1]) sw(0, 1);                   // suqa.cu:536-547
        if (t[1] > t[2]) sw(1, 2);
        if (t[0] > t[1]) sw(0, 1);
        if (t[2] != SUQA_PAULI_Z || t[0] != t[1])
            return fail(SUQA_ERR_UNSUPPORTED, "ERROR: unimplemented pauli TP rotation in the selected configuration");
        if (t[0] < 1) return SUQA_OK;                // suqa_cpu.hpp:493-508: no matching case
        add(t[0], q[0]);
        add(t[1], q[1]);
        add(SUQA_PAULI_Z, q[2]);
    }
    return pauli_masks(c, xm, ym, zm, phase, gc);
}

int suqa_hip_apply_pauli_rotation_general(suqa_hip_ctx* c, const uint32_t* qs, const uint32_t* types, uint32_t n,
                                          double phase, uint32_t gc) {
    SQE_CTX(c);
    if (n > 32 || (n && (!qs || !types))) return fail(SUQA_ERR_ARG, "bad Pauli string");
    int rc = check_qubits(c, qs, n, gc);
    if (rc) return rc;
    uint32_t xm = 0, ym = 0, zm = 0;
    for (uint32_t k = 0; k < n; ++k) {
        if (types[k] > 3u) return fail(SUQA_ERR_ARG, "Pauli type must be 0..3");
        if (types[k] == SUQA_PAULI_X) xm |= 1u << qs[k];
        else if (types[k] == SUQA_PAULI_Y) ym |= 1u << qs[k];
        else if (types[k] == SUQA_PAULI_Z) zm |= 1u << qs[k];
    }
    return pauli_masks(c, xm, ym, zm, phase, gc);
}

int suqa_hip_apply_qft(suqa_hip_ctx* c, const uint32_t* qs, uint32_t n, uint32_t gc) {
    SQE_CTX(c);
    const int qsize = (int)n;                        // suqa.cu:635-647
    for (int o = qsize - 1; o >= 0; o--) {
        int rc = suqa_hip_apply_h(c, qs[o], gc);
        if (rc) return rc;
        for (int in = o - 1; in >= 0; in--) {
            rc = qft_crm(c, qs[in], qs[o], +1 + (o - in), qsize, gc);
            if (rc) return rc;
        }
    }
    return SUQA_OK;
}

int suqa_hip_apply_qft_inverse(suqa_hip_ctx* c, const uint32_t* qs, uint32_t n, uint32_t gc) {
    SQE_CTX(c);
    const int qsize = (int)n;                        // suqa.cu:649-661
    for (int o = 0; o < qsize; o++) {
        for (int in = 0; in < o; in++) {
            int rc = qft_crm(c, qs[in], qs[o], -1 - (o - in), qsize, gc);
            if (rc) return rc;
        }
        int rc = suqa_hip_apply_h(c, qs[o], gc);
        if (rc) return rc;
    }
    return SUQA_OK;
}

// Small states: measurements (and resets) in one single-workgroup launch,
// outcomes copied back only when the caller asks for them.
static int measure_small(suqa_hip_ctx* c, const uint32_t* qs, const double* rds, uint32_t n, bool reset,
                         uint32_t gc, uint32_t* cs) {
    if (cs && c->capturing) return fail(SUQA_ERR_ARG, "a measurement cannot be read back while a graph is being captured");
    if (sqe::tiny_prog(c)) {   // queued beside the gates: one launch for the queue and the measurements
        if (n > sqe::kSlots) return fail(SUQA_ERR_ARG, "too many measurements in one call");
        const uint32_t slot = sqe::alloc_slots(c, n);
        sqe::queue_measure(c, qs, rds, n, reset, gc, slot);
        if (!cs) {
            if (c->queue.size() >= sqe::auto_flush_at(c)) SQE_HIP(sqe::flush_auto(c));
            return SUQA_OK;
        }
        SQE_HIP(sqe::flush(c));
        SQE_HIP(hipMemcpyAsync(c->h_slots, c->d_slots + slot, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, c->stream));
        SQE_HIP(hipStreamSynchronize(c->stream));
        for (uint32_t k = 0; k < n; ++k) cs[k] = c->h_slots[k];
        return SUQA_OK;
    }
    for (uint32_t k0 = 0; k0 < n; k0 += sqe::kMeasMax) {
        const uint32_t m = std::min<uint32_t>(sqe::kMeasMax, n - k0);
        SQE_HIP(sqe::launch_measure_small(c, qs + k0, rds + k0, m, reset ? ~0u : 0u, gc));
        if (cs) {
            SQE_HIP(hipMemcpyAsync(c->h_flags, c->d_flags, sizeof(uint32_t) * m, hipMemcpyDeviceToHost, c->stream));
            SQE_HIP(hipStreamSynchronize(c->stream));
            for (uint32_t k = 0; k < m; ++k) cs[k0 + k] = c->h_flags[k];
        }
    }
    return SUQA_OK;
}

int suqa_hip_measure_qbit(suqa_hip_ctx* c, uint32_t q, double rdoub, uint32_t* cout) {
    SQE_CTX(c);
    if (q >= c->nq) return fail(SUQA_ERR_ARG, "qubit index out of range");
    if (c->nq <= sqe::kMeasSmallBits) {
        uint32_t cc = 0;
        int rc = measure_small(c, &q, &rdoub, 1, false, 0u, &cc);
        if (rc == SUQA_OK && cout) *cout = cc;
        return rc;
    }
    Sel sel;
    sqe::make_sel(c->nq, 0, 0, &sel);
    const uint64_t qb = 1ull << q;
    SQE_HIP(sqe::launch_norm(c, sel, qb, "measure_norm", 16.0 * (double)c->N));
    int rc = read_result(c, 2);
    if (rc) return rc;
    const double p0 = c->h_result[0], p1 = c->h_result[1];
    const uint32_t cc = (uint32_t)(rdoub < p1);      // suqa.cu:690
    // set_ampl_to_zero + vnormalize (suqa.cu:694-695): the norm after the
    // collapse is the kept half's.
    const double vec_norm = std::sqrt(cc ? p1 : p0);
    SQE_HIP(sqe::launch_collapse(c, qb, cc, 1. / vec_norm));
    if (cout) *cout = cc;
    return SUQA_OK;
}

int suqa_hip_apply_reset(suqa_hip_ctx* c, uint32_t q, double rdoub, uint32_t gc) {
    SQE_CTX(c);
    if (q >= c->nq) return fail(SUQA_ERR_ARG, "qubit index out of range");
    if (c->nq <= sqe::kMeasSmallBits) return measure_small(c, &q, &rdoub, 1, true, gc, nullptr);
    uint32_t cc = 0;
    int rc = suqa_hip_measure_qbit(c, q, rdoub, &cc);
    if (rc) return rc;
    if (cc) return suqa_hip_apply_x(c, q, gc);
    return SUQA_OK;
}

int suqa_hip_measure_qbits(suqa_hip_ctx* c, const uint32_t* qs, const double* rdoubs, uint32_t n, uint32_t* cs) {
    SQE_CTX(c);
    if (n && (!qs || !rdoubs)) return fail(SUQA_ERR_ARG, "null argument");
    for (uint32_t k = 0; k < n; ++k)
        if (qs[k] >= c->nq) return fail(SUQA_ERR_ARG, "qubit index out of range");
    if (n && c->nq <= sqe::kMeasSmallBits) return measure_small(c, qs, rdoubs, n, false, 0u, cs);
    for (uint32_t k = 0; k < n; ++k)
        if (int rc = suqa_hip_measure_qbit(c, qs[k], rdoubs[k], cs ? &cs[k] : nullptr)) return rc;
    return SUQA_OK;
}

int suqa_hip_apply_resets(suqa_hip_ctx* c, const uint32_t* qs, const double* rdoubs, uint32_t n, uint32_t gc) {
    SQE_CTX(c);
    if (n && (!qs || !rdoubs)) return fail(SUQA_ERR_ARG, "null argument");
    for (uint32_t k = 0; k < n; ++k)
        if (qs[k] >= c->nq) return fail(SUQA_ERR_ARG, "qubit index out of range");
    if (n && c->nq <= sqe::kMeasSmallBits) return measure_small(c, qs, rdoubs, n, true, gc, nullptr);
    for (uint32_t k = 0; k < n; ++k)
        if (int rc = suqa_hip_apply_reset(c, qs[k], rdoubs[k], gc)) return rc;
    return SUQA_OK;
}

static_assert(SUQA_HIP_OUTCOME_SLOTS == sqe::kSlots, "outcome ring size");

int suqa_hip_measure_qbits_deferred(suqa_hip_ctx* c, const uint32_t* qs, const double* rdoubs, uint32_t n,
                                    uint32_t* slot) {
    SQE_CTX(c);
    if (!slot || (n && (!qs || !rdoubs))) return fail(SUQA_ERR_ARG, "null argument");
    const bool prog = sqe::tiny_prog(c);
    if (!prog && (c->f32 || !c->d_slots || c->nq > sqe::kMeasSmallBits))
        return fail(SUQA_ERR_UNSUPPORTED, "deferred measurements need a state of at most 2^14 amplitudes");
    if (n > sqe::kSlots / 2) return fail(SUQA_ERR_ARG, "too many measurements in one call");
    for (uint32_t k = 0; k < n; ++k)
        if (qs[k] >= c->nq) return fail(SUQA_ERR_ARG, "qubit index out of range");
    *slot = sqe::alloc_slots(c, n);
    if (prog) {   // beside the queued gates
        sqe::queue_measure(c, qs, rdoubs, n, false, 0u, *slot);
        if (c->queue.size() >= sqe::auto_flush_at(c)) SQE_HIP(sqe::flush_auto(c));
        return SUQA_OK;
    }
    if (c->capturing) return fail(SUQA_ERR_ARG, "deferred measurements are not captured in graphs on this state");
    for (uint32_t k0 = 0; k0 < n; k0 += sqe::kMeasMax) {   // at once, the outcomes into the ring
        const uint32_t m = std::min<uint32_t>(sqe::kMeasMax, n - k0);
        SQE_HIP(sqe::launch_measure_small(c, qs + k0, rdoubs + k0, m, 0u, 0u, c->d_slots + *slot + k0));
    }
    return SUQA_OK;
}

int suqa_hip_read_outcomes(suqa_hip_ctx* c, uint32_t slot, uint32_t n, uint32_t* out) {
    SQE_CTX(c);
    if (n && !out) return fail(SUQA_ERR_ARG, "null argument");
    if (!c->d_slots) return fail(SUQA_ERR_UNSUPPORTED, "no outcome slots on this context");
    if (slot >= sqe::kSlots || n > sqe::kSlots) return fail(SUQA_ERR_ARG, "outcome slot range out of the ring");
    if (c->capturing) return fail(SUQA_ERR_ARG, "outcomes cannot be read back while a graph is being captured");
    SQE_HIP(sqe::flush(c));
    const uint32_t n1 = std::min(n, sqe::kSlots - slot);   // up to the ring's end, then from 0
    if (n1) SQE_HIP(hipMemcpyAsync(c->h_slots, c->d_slots + slot, sizeof(uint32_t) * n1, hipMemcpyDeviceToHost, c->stream));
    if (n > n1) SQE_HIP(hipMemcpyAsync(c->h_slots + n1, c->d_slots, sizeof(uint32_t) * (n - n1), hipMemcpyDeviceToHost, c->stream));
    SQE_HIP(hipStreamSynchronize(c->stream));
    for (uint32_t k = 0; k < n; ++k) out[k] = c->h_slots[k];
    return SUQA_OK;
}

int suqa_hip_apply_x_if(suqa_hip_ctx* c, const uint32_t* qs, uint32_t n, uint32_t slot, uint32_t gc) {
    SQE_CTX(c);
    if (n && !qs) return fail(SUQA_ERR_ARG, "null argument");
    const bool prog = sqe::tiny_prog(c);
    if (!prog && (c->f32 || !c->d_slots || c->nq > sqe::kMeasSmallBits))
        return fail(SUQA_ERR_UNSUPPORTED, "conditional gates need a state of at most 2^14 amplitudes");
    if (slot >= sqe::kSlots) return fail(SUQA_ERR_ARG, "outcome slot out of the ring");
    if ((uint64_t)gc >> c->nq) return fail(SUQA_ERR_ARG, "gc_mask has qubits beyond the state");
    for (uint32_t k = 0; k < n; ++k) {
        if (qs[k] >= c->nq) return fail(SUQA_ERR_ARG, "qubit index out of range");
        if ((gc >> qs[k]) & 1u) return fail(SUQA_ERR_ARG, "a conditional X on a qubit of the gc_mask");
    }
    if (!prog) {   // one launch after the queued gates, reading the slots on the device
        if (slot + n > sqe::kSlots) return fail(SUQA_ERR_ARG, "outcome slots run past the ring's end");
        for (uint32_t k0 = 0; k0 < n; k0 += sqe::kMeasMax)
            SQE_HIP(sqe::launch_x_if(c, qs + k0, c->d_slots + slot + k0, std::min<uint32_t>(sqe::kMeasMax, n - k0), gc));
        return SUQA_OK;
    }
    for (uint32_t k = 0; k < n; ++k) {   // pair_1q's selection (i & (gc|T)) == gc|T on i = j|T
        const uint64_t T = 1ull << qs[k];
        sqe::GateDesc d{};
        d.kind = sqe::OP_SWAP;
        d.M = (uint64_t)gc | T;
        d.V = (uint64_t)gc;
        d.X = T;
        d.tag = "swap_pair_if";
        d.alg_bytes = 64.0 * (double)(c->N >> __builtin_popcount(gc | (uint32_t)T));
        d.cond_ref = (slot + k) % sqe::kSlots;
        c->queue.push_back(d);
    }
    if (c->queue.size() >= sqe::auto_flush_at(c)) SQE_HIP(sqe::flush_auto(c));
    return SUQA_OK;
}

int suqa_hip_resets_then_x_if(suqa_hip_ctx* c, const uint32_t* qs, const double* rdoubs, const uint32_t* xqs,
                              uint32_t n, uint32_t slot, uint32_t gc) {
    SQE_CTX(c);
    if (n && (!qs || !rdoubs || !xqs)) return fail(SUQA_ERR_ARG, "null argument");
    if (sqe::tiny_prog(c) || c->f32 || !c->d_slots || c->nq > sqe::kMeasSmallBits) {
        for (uint32_t k = 0; k < n; ++k) {   // queued beside the gates, or refused as the calls it stands for
            if (int rc = suqa_hip_apply_reset(c, qs[k], rdoubs[k], gc)) return rc;
            if (int rc = suqa_hip_apply_x_if(c, xqs + k, 1, (slot + k) % sqe::kSlots, gc)) return rc;
        }
        return SUQA_OK;
    }
    if (slot >= sqe::kSlots || slot + n > sqe::kSlots) return fail(SUQA_ERR_ARG, "outcome slots run past the ring's end");
    if ((uint64_t)gc >> c->nq) return fail(SUQA_ERR_ARG, "gc_mask has qubits beyond the state");
    for (uint32_t k = 0; k < n; ++k) {
        if (qs[k] >= c->nq || xqs[k] >= c->nq) return fail(SUQA_ERR_ARG, "qubit index out of range");
        if ((gc >> xqs[k]) & 1u) return fail(SUQA_ERR_ARG, "a conditional X on a qubit of the gc_mask");
    }
    for (uint32_t k0 = 0; k0 < n; k0 += sqe::kMeasMax) {   // one launch: each reset, then its flip
        const uint32_t m = std::min<uint32_t>(sqe::kMeasMax, n - k0);
        SQE_HIP(sqe::launch_measure_small(c, qs + k0, rdoubs + k0, m, ~0u, gc, nullptr, xqs + k0,
                                          c->d_slots + slot + k0));
    }
    return SUQA_OK;
}

int suqa_hip_prob_filter(suqa_hip_ctx* c, const uint32_t* qs, const uint32_t* q_mask, uint32_t n, double* prob) {
    SQE_CTX(c);
    if (n && (!qs || !q_mask)) return fail(SUQA_ERR_ARG, "null argument");
    uint32_t mqs = 0u, mask = 0u;                    // suqa.cu:706-713
    if (!sqe::prob_filter_masks(qs, q_mask, n, &mqs, &mask)) return fail(SUQA_ERR_ARG, "qubit index out of range");
    Sel sel;
    if (!sqe::make_sel(c->nq, mqs, mask, &sel)) {
        if (prob) *prob = 0.0;
        return SUQA_OK;
    }
    SQE_HIP(sqe::launch_norm(c, sel, 0, "prob_filter", 16.0 * (double)sel.count));
    int rc = read_result(c, 1);
    if (rc) return rc;
    if (prob) *prob = c->h_result[0];
    return SUQA_OK;
}

namespace {

// fs1/fs2 exactly as qms.cuh:291-300 evaluates them (host libm).
void w_coef(uint64_t En, uint64_t Eo, double cf, double* out) {
    const double dE = ((double)(uint32_t)En * cf - (double)(uint32_t)Eo * cf);
    if (dE > 0) {
        out[0] = std::exp(-dE / 2.0);
        out[1] = std::sqrt(1.0 - out[0] * out[0]);
    } else {
        out[0] = 1.0;
        out[1] = 0.0;
    }
}

// The (E_old, E_new) coefficient table of an L-level register in c->d_wtab
// (L <= 256), rebuilt when c_factor, the mask or the offset change.
int w_table(suqa_hip_ctx* c, uint32_t wmask, uint32_t b0, double cf, uint64_t L) {
    if (c->wtab_valid && c->wtab_c == cf && c->wtab_mask == wmask && c->wtab_b0 == b0) return SUQA_OK;
    SQE_HIP(sqe::flush(c));   // a queued apply_W reads the table it was queued with
    std::vector<double> tab(2 * L * L);
    for (uint64_t eo = 0; eo < L; ++eo)
        for (uint64_t en = 0; en < L; ++en) w_coef(en, eo, cf, &tab[2 * (eo * L + en)]);
    if (c->wtab_L < L) {
        if (c->d_wtab) SQE_HIP(hipFree(c->d_wtab));
        c->d_wtab = nullptr;
        void* p = nullptr;
        SQE_HIP(hipMalloc(&p, sizeof(double) * 2 * 256 * 256));
        c->d_wtab = static_cast<double*>(p);
        c->wtab_L = 256;
    }
    SQE_HIP(hipMemcpyAsync(c->d_wtab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, c->stream));
    SQE_HIP(hipStreamSynchronize(c->stream));
    c->wtab_valid = true;
    c->wtab_c = cf;
    c->wtab_mask = wmask;
    c->wtab_b0 = b0;
    return SUQA_OK;
}

// An energy register mask: contiguous bits from b0.
bool reg_mask_ok(uint32_t mask, uint32_t b0) {
    if (b0 >= 32 || mask == 0) return false;
    const uint32_t m = mask >> b0;
    return (mask & ((1u << b0) - 1u)) == 0 && (m & (m + 1u)) == 0;
}

}  // namespace

int suqa_hip_qms_apply_W(suqa_hip_ctx* c, uint32_t q_acc, uint32_t wmask, uint32_t b0, double cf,
                         uint32_t Eold) {
    SQE_CTX(c);
    if (q_acc != c->nq - 1) return fail(SUQA_ERR_ARG, "apply_W needs the acceptance qubit to be the most significant");
    if (b0 >= 32) return fail(SUQA_ERR_ARG, "bad energy register offset");
    const uint64_t L = (uint64_t)(wmask >> b0) + 1ull;
    if (L > (1ull << 20)) return fail(SUQA_ERR_ARG, "energy register too wide for apply_W");
    const double* row = nullptr;
    if (L <= 256 && Eold < L) {
        if (int rc = w_table(c, wmask, b0, cf, L)) return rc;
        row = c->d_wtab + 2 * (uint64_t)Eold * L;
    } else {
        std::vector<double> r(2 * L);
        for (uint64_t en = 0; en < L; ++en) w_coef(en, Eold, cf, &r[2 * en]);
        SQE_HIP(sqe::flush(c));   // a queued apply_W reads the row it was queued with
        if (c->wrow_cap < L) {
            if (c->d_wrow) SQE_HIP(hipFree(c->d_wrow));
            c->d_wrow = nullptr;
            void* p = nullptr;
            SQE_HIP(hipMalloc(&p, sizeof(double) * 2 * L));
            c->d_wrow = static_cast<double*>(p);
            c->wrow_cap = L;
        }
        SQE_HIP(hipMemcpyAsync(c->d_wrow, r.data(), sizeof(double) * r.size(), hipMemcpyHostToDevice, c->stream));
        SQE_HIP(hipStreamSynchronize(c->stream));
        row = c->d_wrow;
    }
    SQE_HIP(sqe::launch_qms_W(c, row, wmask, b0, 0u, 0u, (uint32_t)L));
    return SUQA_OK;
}

int suqa_hip_qms_apply_W_deferred(suqa_hip_ctx* c, uint32_t q_acc, uint32_t wmask, uint32_t b0, double cf,
                                  uint32_t eold_slot) {
    SQE_CTX(c);
    if (!sqe::tiny_prog(c)) return fail(SUQA_ERR_UNSUPPORTED, "deferred apply_W needs a state of one small pass tile");
    if (q_acc != c->nq - 1) return fail(SUQA_ERR_ARG, "apply_W needs the acceptance qubit to be the most significant");
    if (!reg_mask_ok(wmask, b0) || (wmask >> q_acc) != 0) return fail(SUQA_ERR_ARG, "bad energy register");
    if (eold_slot >= sqe::kSlots) return fail(SUQA_ERR_ARG, "outcome slot out of the ring");
    const uint64_t L = (uint64_t)(wmask >> b0) + 1ull;
    if (int rc = w_table(c, wmask, b0, cf, L)) return rc;
    sqe::GateDesc d{};
    d.kind = sqe::kWOp;
    d.tab = c->d_wtab;
    d.wmask = wmask;
    d.b0 = b0;
    d.L = (uint32_t)L;
    d.eold_ref = eold_slot;
    d.eold_bits = (uint32_t)__builtin_popcount(wmask);
    d.tag = "qms_W";
    c->queue.push_back(d);
    if (c->queue.size() >= sqe::auto_flush_at(c)) SQE_HIP(sqe::flush_auto(c));
    return SUQA_OK;
}

int suqa_hip_qms_apply_W2(suqa_hip_ctx* c, uint32_t q_acc, uint32_t W_mask_Enew, uint32_t bm_enes_new0,
                          uint32_t W_mask_Eold, uint32_t bm_enes_old0, double cf) {
    SQE_CTX(c);
    if (q_acc != c->nq - 1) return fail(SUQA_ERR_ARG, "apply_W needs the acceptance qubit to be the most significant");
    if (!reg_mask_ok(W_mask_Enew, bm_enes_new0) || !reg_mask_ok(W_mask_Eold, bm_enes_old0))
        return fail(SUQA_ERR_ARG, "energy registers must be contiguous bit ranges at their offsets");
    if ((W_mask_Enew & W_mask_Eold) != 0 || ((W_mask_Enew | W_mask_Eold) >> q_acc) != 0)
        return fail(SUQA_ERR_ARG, "energy registers overlap each other or the acceptance qubit");
    const uint64_t L = (uint64_t)(W_mask_Enew >> bm_enes_new0) + 1ull;
    if ((uint64_t)(W_mask_Eold >> bm_enes_old0) + 1ull != L) return fail(SUQA_ERR_ARG, "energy registers differ in width");
    if (L > 256) return fail(SUQA_ERR_ARG, "energy registers too wide for the two-register apply_W (at most 8 qubits)");
    if (int rc = w_table(c, W_mask_Enew, bm_enes_new0, cf, L)) return rc;
    SQE_HIP(sqe::launch_qms_W(c, c->d_wtab, W_mask_Enew, bm_enes_new0, W_mask_Eold, bm_enes_old0, (uint32_t)L));
    return SUQA_OK;
}

int suqa_hip_state_info(suqa_hip_ctx* c, uint32_t* nq, uint64_t* N, double** re, double** im) {
    SQE_CTX(c);
    SQE_HIP(sqe::flush(c));
    if (nq) *nq = c->nq;
    if (N) *N = c->N;
    if (re) *re = c->d_re;
    if (im) *im = c->d_im;
    return SUQA_OK;
}

int suqa_hip_download(suqa_hip_ctx* c, double* re, double* im) {
    SQE_CTX(c);
    SQE_HIP(sqe::flush(c));
    if (re) SQE_HIP(hipMemcpyAsync(re, c->d_re, c->N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (im) SQE_HIP(hipMemcpyAsync(im, c->d_im, c->N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SQE_HIP(hipStreamSynchronize(c->stream));
    return SUQA_OK;
}

int suqa_hip_upload(suqa_hip_ctx* c, const double* re, const double* im) {
    SQE_CTX(c);
    SQE_HIP(sqe::flush(c));
    if (re) SQE_HIP(hipMemcpyAsync(c->d_re, re, c->N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (im) SQE_HIP(hipMemcpyAsync(c->d_im, im, c->N * sizeof(double), hipMemcpyHostToDevice, c->stream));
    SQE_HIP(hipStreamSynchronize(c->stream));
    return SUQA_OK;
}

int suqa_hip_synchronize(suqa_hip_ctx* c) {
    SQE_CTX(c);
    SQE_HIP(sqe::flush(c));
    SQE_HIP(hipStreamSynchronize(c->stream));
    return SUQA_OK;
}

int suqa_hip_set_tile_bits(suqa_hip_ctx* c, uint32_t bits) {
    SQE_CTX(c);
    if (bits != 0 && (bits < sqe::kMinTileBits || bits > sqe::kMaxTileBits))
        return fail(SUQA_ERR_ARG, "tile bits must be 0 (one sweep per gate) or 1..12");
    SQE_HIP(sqe::flush(c));
    c->tile_bits = bits;
    return SUQA_OK;
}

int suqa_hip_set_gate_order(suqa_hip_ctx* c, int packed) {
    SQE_CTX(c);
    if (packed != 0 && packed != 1) return fail(SUQA_ERR_ARG, "gate order must be 0 (program order) or 1 (packed)");
    SQE_HIP(sqe::flush(c));   // gates queued so far keep the order they were queued in
    c->gate_reorder = packed == 1;
    return SUQA_OK;
}

int suqa_hip_autotune_tile_bits(suqa_hip_ctx* c, const suqa_hip_gate* gates, uint32_t ngates,
                                const uint32_t* candidates, uint32_t ncand, uint32_t* best, double* ms_per_gate) {
    SQE_CTX(c);
    if (!gates || !ngates || !candidates || !ncand || !best) return fail(SUQA_ERR_ARG, "empty sample or candidates");
    for (uint32_t k = 0; k < ncand; ++k)
        if (candidates[k] > sqe::kMaxTileBits) return fail(SUQA_ERR_ARG, "tile bits must be 0..12");
    const uint32_t keep = c->tile_bits;
    double best_ms = 1e300;
    *best = keep;
    for (uint32_t k = 0; k < ncand; ++k) {
        int rc = suqa_hip_set_tile_bits(c, candidates[k]);
        if (rc) return rc;
        // warm-up on a few gates, then the timed sample
        rc = suqa_hip_run_gates(c, gates, std::min<uint32_t>(ngates, 8));
        if (!rc) rc = suqa_hip_synchronize(c);
        if (rc) return rc;
        const auto t0 = std::chrono::steady_clock::now();
        rc = suqa_hip_run_gates(c, gates, ngates);
        if (!rc) rc = suqa_hip_synchronize(c);
        if (rc) return rc;
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (ms_per_gate) ms_per_gate[k] = ms / ngates;
        if (ms < best_ms) {
            best_ms = ms;
            *best = candidates[k];
        }
    }
    return suqa_hip_set_tile_bits(c, *best);
}

int suqa_hip_get_tile_bits(suqa_hip_ctx* c, uint32_t* bits) {
    SQE_CTX(c);
    if (bits) *bits = c->tile_bits;
    return SUQA_OK;
}

int suqa_hip_pass_stats(suqa_hip_ctx* c, uint64_t* passes, uint64_t* pass_gates, uint64_t* single_sweeps) {
    SQE_CTX(c);
    SQE_HIP(sqe::flush(c));
    if (passes) *passes = c->passes;
    if (pass_gates) *pass_gates = c->pass_gates;
    if (single_sweeps) *single_sweeps = c->single_sweeps;
    return SUQA_OK;
}

int suqa_hip_run_gates(suqa_hip_ctx* c, const suqa_hip_gate* gates, uint32_t ngates) {
    SQE_CTX(c);
    const uint32_t tx = SUQA_PAULI_X, tz = SUQA_PAULI_Z;
    for (uint32_t k = 0; k < ngates; ++k) {
        const suqa_hip_gate& g = gates[k];
        int rc = SUQA_OK;
        switch (g.kind) {
        case SUQA_GATE_H: rc = suqa_hip_apply_h(c, g.target, 0); break;
        case SUQA_GATE_X: rc = suqa_hip_apply_x(c, g.target, 0); break;
        case SUQA_GATE_RX: rc = suqa_hip_apply_pauli_TP_rotation(c, &g.target, &tx, 1, -g.angle / 2.0, 0); break;
        case SUQA_GATE_RZ: rc = suqa_hip_apply_pauli_TP_rotation(c, &g.target, &tz, 1, -g.angle / 2.0, 0); break;
        case SUQA_GATE_U1: rc = suqa_hip_apply_u1(c, g.target, 1u, g.angle, 0); break;
        case SUQA_GATE_CX:
        case SUQA_GATE_MCX:
            if (g.nctrl > 3) return fail(SUQA_ERR_ARG, "at most 3 controls per workload gate");
            rc = suqa_hip_apply_mcx(c, g.ctrl, nullptr, g.nctrl, g.target, 0);
            break;
        default: return fail(SUQA_ERR_ARG, "unknown gate kind");
        }
        if (rc) return rc;
    }
    return SUQA_OK;
}

int suqa_hip_run_paulis(suqa_hip_ctx* c, const suqa_hip_pauli_op* ops, uint32_t nops) {
    SQE_CTX(c);
    if (nops && !ops) return fail(SUQA_ERR_ARG, "null op list");
    for (uint32_t k = 0; k < nops; ++k)
        if (ops[k].w > 8) return fail(SUQA_ERR_ARG, "Pauli op weight above 8");
    std::vector<uint32_t> order;
    if (c->tile_bits > 0 && c->pauli_reorder) order = sqe::pauli_pass_order(c, ops, nops);
    for (uint32_t k = 0; k < nops; ++k) {
        const suqa_hip_pauli_op& o = ops[order.empty() ? k : order[k]];
        int rc = suqa_hip_apply_pauli_rotation_general(c, o.qs, o.types, o.w, o.phase, o.gc_mask);
        if (rc) return rc;
    }
    return SUQA_OK;
}



int suqa_hip_timing_enable(suqa_hip_ctx* c, int on) {
    SQE_CTX(c);
    if (!on) sqe::timing_flush(c);
    c->timing = on != 0;
    return SUQA_OK;
}

int suqa_hip_timing_reset(suqa_hip_ctx* c) {
    SQE_CTX(c);
    sqe::timing_flush(c);
    c->stats.clear();
    return SUQA_OK;
}

int suqa_hip_timing_report(suqa_hip_ctx* c, suqa_hip_kstat* out, uint32_t max, uint32_t* n) {
    SQE_CTX(c);
    SQE_HIP(sqe::flush(c));
    sqe::timing_flush(c);
    uint32_t k = 0;
    for (; k < c->stats.size() && k < max; ++k) {
        std::memset(out[k].name, 0, sizeof(out[k].name));
        std::strncpy(out[k].name, c->stats[k].name.c_str(), sizeof(out[k].name) - 1);
        out[k].launches = c->stats[k].launches;
        out[k].total_ms = c->stats[k].total_ms;
        out[k].alg_bytes = c->stats[k].alg_bytes;
        out[k].flops = 0.0;
    }
    if (n) *n = (uint32_t)c->stats.size();
    return SUQA_OK;
}

int suqa_hip_graph_begin(suqa_hip_ctx* c) {
    SQE_CTX(c);
    if (c->capturing) return fail(SUQA_ERR_ARG, "a graph capture is already open on this context");
    SQE_HIP(sqe::flush(c));   // queued gates run before the captured sequence, not in it
    SQE_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    return SUQA_OK;
}

int suqa_hip_graph_end(suqa_hip_ctx* c, uint32_t* graph) {
    SQE_CTX(c);
    if (!graph) return fail(SUQA_ERR_ARG, "null argument");
    if (!c->capturing) return fail(SUQA_ERR_ARG, "no graph capture is open on this context");
    const hipError_t ef = sqe::flush(c);   // the captured gates' passes
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(c->stream, &g);
    c->capturing = false;
    if (ef != hipSuccess || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return fail(SUQA_ERR_HIP, std::string("graph capture failed: ") + hipGetErrorString(ef != hipSuccess ? ef : ee));
    }
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) return fail(SUQA_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    c->graphs.push_back(x);
    *graph = (uint32_t)(c->graphs.size() - 1);
    return SUQA_OK;
}

int suqa_hip_graph_launch(suqa_hip_ctx* c, uint32_t graph) {
    SQE_CTX(c);
    if (c->capturing) return fail(SUQA_ERR_ARG, "a graph cannot be launched while one is being captured");
    if (graph >= c->graphs.size() || !c->graphs[graph]) return fail(SUQA_ERR_ARG, "unknown graph");
    SQE_HIP(sqe::flush(c));   // gates queued before it run first
    SQE_HIP(hipGraphLaunch(c->graphs[graph], c->stream));
    return SUQA_OK;
}

int suqa_hip_copy_kernel(suqa_hip_ctx* c, void* dst, const void* src, uint64_t nbytes) {
    SQE_CTX(c);
    if ((nbytes & 15u) || ((uintptr_t)dst & 15u) || ((uintptr_t)src & 15u))
        return fail(SUQA_ERR_ARG, "copy needs 16-byte aligned pointers and size");
    SQE_HIP(sqe::launch_copy(c, dst, src, nbytes));
    return SUQA_OK;
}

int suqa_hip_device_alloc(uint64_t nbytes, void** ptr) {
    if (!ptr) return fail(SUQA_ERR_ARG, "null pointer");
    if (hipMalloc(ptr, nbytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(SUQA_ERR_NOMEM, "hipMalloc failed");
    }
    return SUQA_OK;
}

int suqa_hip_device_free(void* ptr) {
    SQE_HIP(hipFree(ptr));
    return SUQA_OK;
}

}  // extern "C"
