// fuse.cpp -- host side of the precision pattern: gate matrices and greedy
// fusion of a gate list into dense K-qubit blocks (computed in FP64).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "fused.hpp"

namespace sqf {

using cd = std::complex<double>;

// Workload gate -> (sorted qubits, 2^g x 2^g matrix in that qubit order).
// H uses the reference's 1/sqrt(2) constant (suqa.cuh:57); RX(a) and RZ(a)
// are exp(-i a P / 2) as apply_pauli_TP_rotation({q},{P},-a/2) computes them.
void gate_matrix(const suqa_hip_gate& g, std::vector<uint32_t>& qs, std::vector<cd>& m) {
    const double T = 0.7071067811865475;
    qs.clear();
    if (g.kind == SUQA_GATE_CX || g.kind == SUQA_GATE_MCX) {
        for (uint32_t k = 0; k < g.nctrl; ++k) qs.push_back(g.ctrl[k]);
        qs.push_back(g.target);
        std::sort(qs.begin(), qs.end());
        const uint32_t d = 1u << qs.size();
        uint32_t tb = 0, cm = 0;
        for (uint32_t j = 0; j < qs.size(); ++j) {
            if (qs[j] == g.target) tb = 1u << j;
            else cm |= 1u << j;
        }
        m.assign((size_t)d * d, cd(0, 0));
        for (uint32_t in = 0; in < d; ++in) {
            const uint32_t out = ((in & cm) == cm) ? (in ^ tb) : in;
            m[(size_t)out * d + in] = 1.0;
        }
        return;
    }
    qs.push_back(g.target);
    m.assign(4, cd(0, 0));
    switch (g.kind) {
    case SUQA_GATE_H: m = {T, T, T, -T}; break;
    case SUQA_GATE_X: m = {0, 1, 1, 0}; break;
    case SUQA_GATE_RX: {   // exp(i phi X), phi = -a/2: [[c, i s], [i s, c]]
        const double ph = -g.angle / 2.0;
        m = {std::cos(ph), cd(0, std::sin(ph)), cd(0, std::sin(ph)), std::cos(ph)};
        break;
    }
    case SUQA_GATE_RZ: {   // exp(i phi Z): diag(e^{i phi}, e^{-i phi})
        const double ph = -g.angle / 2.0;
        m = {std::polar(1.0, ph), 0, 0, std::polar(1.0, -ph)};
        break;
    }
    case SUQA_GATE_U1: m = {1, 0, 0, std::polar(1.0, g.angle)}; break;
    default: break;
    }
}

namespace {

// Applies gate matrix gm (on block positions pos[]) to every column of the
// block matrix B (D x D, row-major): B <- G B.
void apply_to_block(std::vector<cd>& B, uint32_t K, const std::vector<uint32_t>& pos, const std::vector<cd>& gm) {
    const uint32_t D = 1u << K, gsz = (uint32_t)pos.size(), gd = 1u << gsz;
    uint32_t gmask = 0;
    for (uint32_t p : pos) gmask |= 1u << p;
    std::vector<uint32_t> off(gd);
    for (uint32_t x = 0; x < gd; ++x) {
        uint32_t o = 0;
        for (uint32_t j = 0; j < gsz; ++j)
            if ((x >> j) & 1u) o |= 1u << pos[j];
        off[x] = o;
    }
    std::vector<cd> v(gd), w(gd);
    for (uint32_t col = 0; col < D; ++col) {
        for (uint32_t r = 0; r < D; ++r) {
            if (r & gmask) continue;   // r: the rest index with gate bits 0
            for (uint32_t x = 0; x < gd; ++x) v[x] = B[(size_t)(r | off[x]) * D + col];
            for (uint32_t y = 0; y < gd; ++y) {
                cd s = 0;
                for (uint32_t x = 0; x < gd; ++x) s += gm[(size_t)y * gd + x] * v[x];
                w[y] = s;
            }
            for (uint32_t y = 0; y < gd; ++y) B[(size_t)(r | off[y]) * D + col] = w[y];
        }
    }
}

}  // namespace

std::vector<Block> fuse(const suqa_hip_gate* g, uint32_t n, uint32_t nq, uint32_t width, uint32_t window) {
    std::vector<Block> out;
    std::vector<char> used(n, 0);
    std::vector<std::vector<uint32_t>> gq(n);
    std::vector<std::vector<cd>> gm(n);
    for (uint32_t i = 0; i < n; ++i) gate_matrix(g[i], gq[i], gm[i]);
    uint32_t first = 0;
    while (first < n) {
        while (first < n && used[first]) ++first;
        if (first >= n) break;
        uint64_t S = 0, blocked = 0;
        std::vector<uint32_t> members;
        uint32_t scanned = 0;
        for (uint32_t i = first; i < n && scanned < window; ++i) {
            if (used[i]) continue;
            ++scanned;
            uint64_t G = 0;
            for (uint32_t q : gq[i]) G |= 1ull << q;
            if ((G & blocked) == 0 && (uint32_t)__builtin_popcountll(S | G) <= width) {
                S |= G;
                members.push_back(i);
                used[i] = 1;
            } else {
                blocked |= G;
            }
            // nothing further can join once every qubit is blocked, or the block is
            // full and all of its qubits are blocked
            const bool full = (uint32_t)__builtin_popcountll(S) >= width;
            if ((uint32_t)__builtin_popcountll(blocked) >= nq || (full && (blocked & S) == S)) break;
        }
        // pad the block to exactly `width` qubits with the lowest free qubits
        for (uint32_t q = 0; q < nq && (uint32_t)__builtin_popcountll(S) < width; ++q) S |= 1ull << q;
        Block b;
        b.K = 0;
        for (uint32_t q = 0; q < nq; ++q)
            if ((S >> q) & 1ull) b.qubits[b.K++] = q;
        const uint32_t D = 1u << b.K;
        b.m.assign((size_t)D * D, cd(0, 0));
        for (uint32_t r = 0; r < D; ++r) b.m[(size_t)r * D + r] = 1.0;
        for (uint32_t i : members) {
            std::vector<uint32_t> pos;
            for (uint32_t q : gq[i])
                for (uint32_t j = 0; j < b.K; ++j)
                    if (b.qubits[j] == q) pos.push_back(j);
            apply_to_block(b.m, b.K, pos, gm[i]);
        }
        b.ngates = (uint32_t)members.size();
        out.push_back(std::move(b));
    }
    return out;
}

}  // namespace sqf

extern "C" int suqa_hip_fuse_plan(const suqa_hip_gate* gates, uint32_t ngates, uint32_t nq, uint32_t width,
                                  uint32_t max_blocks, uint32_t* nblocks, uint32_t* qubits, uint32_t* gates_in,
                                  double* mats) {
    if ((ngates && !gates) || !nblocks || width < 1 || width > 6 || nq < width || nq > 64) return SUQA_ERR_ARG;
    const auto blocks = sqf::fuse(gates, ngates, nq, width, 256);
    *nblocks = (uint32_t)blocks.size();
    const size_t msz = (size_t)1 << (2 * width);
    for (uint32_t b = 0; b < blocks.size() && b < max_blocks; ++b) {
        if (qubits)
            for (uint32_t j = 0; j < 6; ++j) qubits[6 * b + j] = j < blocks[b].K ? blocks[b].qubits[j] : 0u;
        if (gates_in) gates_in[b] = blocks[b].ngates;
        if (mats)
            for (size_t k = 0; k < msz; ++k) {
                mats[2 * (b * msz + k)] = blocks[b].m[k].real();
                mats[2 * (b * msz + k) + 1] = blocks[b].m[k].imag();
            }
    }
    return SUQA_OK;
}
