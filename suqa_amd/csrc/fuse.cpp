// fuse.cpp -- host side of the precision pattern: operation matrices and
// greedy fusion of an operation list into dense K-qubit blocks (in FP64).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "fused.hpp"

namespace sqf {

namespace {

const double kT = 0.7071067811865475;   // the reference's 1/sqrt(2) (suqa.cuh:57)

// Builds an op on the qubit set `qs` (any order, distinct) whose matrix is
// f(in) applied column by column: f writes column `in` (indices over the
// sorted qubits) into col[].
template <class F>
Op make_op(std::vector<uint32_t> qs, F f) {
    std::sort(qs.begin(), qs.end());
    Op o;
    o.n = (uint32_t)qs.size();
    for (uint32_t j = 0; j < o.n; ++j) o.qs[j] = qs[j];
    const uint32_t d = 1u << o.n;
    o.m.assign((size_t)d * d, cd(0, 0));
    std::vector<cd> col(d);
    for (uint32_t in = 0; in < d; ++in) {
        std::fill(col.begin(), col.end(), cd(0, 0));
        f(o, in, col);
        for (uint32_t out = 0; out < d; ++out) o.m[(size_t)out * d + in] = col[out];
    }
    return o;
}

uint32_t bit_of(const Op& o, uint32_t q) {
    for (uint32_t j = 0; j < o.n; ++j)
        if (o.qs[j] == q) return 1u << j;
    return 0;
}

// Applies op matrix gm (on block positions pos[]) to every column of the
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
            if (r & gmask) continue;   // r: the rest index with the op's bits 0
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

// RX(a) and RZ(a) are exp(-i a P / 2) as apply_pauli_TP_rotation({q},{P},-a/2)
// computes them; MCX flips the target when every control is 1.
Op gate_op(const suqa_hip_gate& g) {
    if (g.kind == SUQA_GATE_CX || g.kind == SUQA_GATE_MCX) {
        std::vector<uint32_t> qs(g.ctrl, g.ctrl + g.nctrl);
        qs.push_back(g.target);
        return make_op(qs, [&](const Op& o, uint32_t in, std::vector<cd>& col) {
            const uint32_t tb = bit_of(o, g.target);
            uint32_t cm = 0;
            for (uint32_t k = 0; k < g.nctrl; ++k) cm |= bit_of(o, g.ctrl[k]);
            col[((in & cm) == cm) ? (in ^ tb) : in] = 1.0;
        });
    }
    cd m[4] = {1, 0, 0, 1};
    const double ph = -g.angle / 2.0;
    switch (g.kind) {
    case SUQA_GATE_H: m[0] = kT; m[1] = kT; m[2] = kT; m[3] = -kT; break;
    case SUQA_GATE_X: m[0] = 0; m[1] = 1; m[2] = 1; m[3] = 0; break;
    case SUQA_GATE_RX:   // exp(i ph X) = [[c, i s], [i s, c]]
        m[0] = std::cos(ph); m[1] = cd(0, std::sin(ph)); m[2] = cd(0, std::sin(ph)); m[3] = std::cos(ph);
        break;
    case SUQA_GATE_RZ: m[0] = std::polar(1.0, ph); m[1] = 0; m[2] = 0; m[3] = std::polar(1.0, -ph); break;
    case SUQA_GATE_U1: m[3] = std::polar(1.0, g.angle); break;
    default: break;
    }
    Op o;
    o.n = 1;
    o.qs[0] = g.target;
    o.m.assign(m, m + 4);
    return o;
}

Op cx_op(uint64_t ctrl, uint32_t tgt) {
    std::vector<uint32_t> qs;
    for (uint32_t q = 0; q < 64; ++q)
        if ((ctrl >> q) & 1ull) qs.push_back(q);
    qs.push_back(tgt);
    return make_op(qs, [&](const Op& o, uint32_t in, std::vector<cd>& col) {
        const uint32_t tb = bit_of(o, tgt);
        const uint32_t cm = ((1u << o.n) - 1u) & ~tb;
        col[((in & cm) == cm) ? (in ^ tb) : in] = 1.0;
    });
}

PassOp pass_gate(const suqa_hip_gate& g) {
    PassOp p;
    p.op = gate_op(g);
    if (g.kind == SUQA_GATE_CX || g.kind == SUQA_GATE_MCX) {
        for (uint32_t k = 0; k < g.nctrl; ++k) p.ctrl |= 1ull << g.ctrl[k];
        p.tgt = g.target;
    } else if (g.kind == SUQA_GATE_RZ || g.kind == SUQA_GATE_U1) {
        p.diag = true;
        p.dq = g.target;
        p.d0 = p.op.m[0];
        p.d1 = p.op.m[3];
    }
    return p;
}

// exp(i phi P) = cos(phi) I + i sin(phi) P with P|x> = prod_k f_k |x ^ flips>,
// f = 1 (X), i (-1)^b (Y), (-1)^b (Z), on the string's non-identity qubits;
// the gc_mask qubits become the op's controls.
bool pauli_op(const suqa_hip_pauli_op& p, Op& out) {
    if (p.w > 8) return false;
    std::vector<uint32_t> qs;
    uint64_t seen = 0;
    for (uint32_t k = 0; k < p.w; ++k) {
        if (p.types[k] > 3u || p.qs[k] >= 32 || ((seen >> p.qs[k]) & 1ull)) return false;
        seen |= 1ull << p.qs[k];
        if (p.types[k] != 0u) qs.push_back(p.qs[k]);
    }
    if (seen & (uint64_t)p.gc_mask) return false;
    if (qs.size() > (size_t)kMaxK) return false;
    const double c = std::cos(p.phase), s = std::sin(p.phase);
    out = make_op(qs, [&](const Op& o, uint32_t in, std::vector<cd>& col) {
        uint32_t flip = 0;
        cd f = 1.0;
        for (uint32_t k = 0; k < p.w; ++k) {
            const uint32_t b = bit_of(o, p.qs[k]);
            const bool one = (in & b) != 0;
            switch (p.types[k]) {
            case SUQA_PAULI_X: flip |= b; break;
            case SUQA_PAULI_Y: flip |= b; f *= cd(0, one ? -1.0 : 1.0); break;
            case SUQA_PAULI_Z: if (one) f = -f; break;
            default: break;
            }
        }
        col[in] += c;
        col[in ^ flip] += cd(0, s) * f;
    });
    out.cmask = p.gc_mask;
    return true;
}

void pauli_rot(const suqa_hip_pauli_op& p, Op& out) {
    // P|x> = i^ny (-1)^{|x & (Y|Z)|} |x ^ (X|Y)>, so
    // exp(i phi P) psi (y) = cos phi psi(y) + i sin phi i^ny (-1)^{|(y ^ flips) & (Y|Z)|} psi(y ^ flips)
    uint32_t ny = 0;
    out.rflip = out.rpar = 0;
    for (uint32_t k = 0; k < p.w; ++k) {
        const uint64_t b = 1ull << p.qs[k];
        if (p.types[k] == SUQA_PAULI_X || p.types[k] == SUQA_PAULI_Y) out.rflip |= b;
        if (p.types[k] == SUQA_PAULI_Y || p.types[k] == SUQA_PAULI_Z) out.rpar |= b;
        if (p.types[k] == SUQA_PAULI_Y) ++ny;
    }
    const double s = std::sin(p.phase);
    static const cd ipow[4] = {cd(0, 1), cd(-1, 0), cd(0, -1), cd(1, 0)};   // i^(ny + 1)
    out.rot = true;
    out.rc = std::cos(p.phase);
    out.rk = s * ipow[ny & 3u];
    out.m.clear();
}

Op h_op(uint32_t q) {
    suqa_hip_gate g{};
    g.kind = SUQA_GATE_H;
    g.target = q;
    return gate_op(g);
}

Op cphase_op(uint32_t a, uint32_t b, double phase) {
    return make_op({a, b}, [&](const Op&, uint32_t in, std::vector<cd>& col) {
        col[in] = in == 3u ? std::polar(1.0, phase) : cd(1.0, 0.0);
    });
}

// suqa.cu:649-661 order: for o = 0..n-1, the crm(-(1+o-in)) rotations with
// in < o, then H on qs[o]; crm(m) is the controlled phase sign(m) 2 pi / 2^|m|.
void qft_inverse_ops(const uint32_t* qs, uint32_t n, std::vector<Op>& out) {
    for (uint32_t o = 0; o < n; ++o) {
        for (uint32_t in = 0; in < o; ++in) {
            const uint32_t m = 1 + (o - in);
            out.push_back(cphase_op(qs[in], qs[o], -2.0 * M_PI / std::ldexp(1.0, (int)m)));
        }
        out.push_back(h_op(qs[o]));
    }
}

std::vector<Block> fuse(const std::vector<Op>& ops, uint32_t nq, uint32_t width, uint32_t window,
                        uint64_t pad_allowed) {
    const uint32_t n = (uint32_t)ops.size();
    std::vector<Block> out;
    std::vector<char> used(n, 0);
    std::vector<uint64_t> gmask(n, 0);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < ops[i].n; ++j) gmask[i] |= 1ull << ops[i].qs[j];
    uint32_t first = 0;
    while (first < n) {
        while (first < n && used[first]) ++first;
        if (first >= n) break;
        if (ops[first].swap || ops[first].rot) {   // a controlled X or a sparse rotation: its own block
            Block b;
            b.K = 1;
            b.qubits[0] = ops[first].qs[0];
            b.cmask = ops[first].cmask;
            b.swap = ops[first].swap;
            b.rot = ops[first].rot ? &ops[first] : nullptr;
            b.ngates = 1;
            used[first] = 1;
            out.push_back(std::move(b));
            continue;
        }
        const uint64_t cm = ops[first].cmask;
        uint64_t S = 0, lt = 0, lc = 0;   // block targets; targets / controls of ops left out
        std::vector<uint32_t> members;
        uint32_t scanned = 0;
        for (uint32_t i = first; i < n && scanned < window; ++i) {
            if (used[i]) continue;
            ++scanned;
            const uint64_t G = gmask[i], C = ops[i].cmask;
            const bool commutes = (G & (lt | lc)) == 0 && (C & lt) == 0;
            if (!ops[i].swap && !ops[i].rot && commutes && C == cm && (uint32_t)__builtin_popcountll(S | G) <= width) {
                S |= G;
                members.push_back(i);
                used[i] = 1;
            } else {
                lt |= G;
                lc |= C;
            }
            // nothing further can join once every qubit is blocked, the block's
            // control is some left-out op's target, or the block is full and all
            // of its qubits are blocked
            const bool full = (uint32_t)__builtin_popcountll(S) >= width;
            if ((uint32_t)__builtin_popcountll(lt | lc) >= nq || (lt & cm) != 0 || (full && ((lt | lc) & S) == S))
                break;
        }
        // pad the block to exactly `width` qubits with the lowest free allowed non-control qubits
        for (uint32_t q = 0; q < nq && (uint32_t)__builtin_popcountll(S) < width; ++q)
            if (!((cm >> q) & 1ull) && ((pad_allowed >> q) & 1ull)) S |= 1ull << q;
        Block b;
        b.K = 0;
        b.cmask = cm;
        for (uint32_t q = 0; q < nq; ++q)
            if ((S >> q) & 1ull) b.qubits[b.K++] = q;
        const uint32_t D = 1u << b.K;
        b.m.assign((size_t)D * D, cd(0, 0));
        for (uint32_t r = 0; r < D; ++r) b.m[(size_t)r * D + r] = 1.0;
        for (uint32_t i : members) {
            std::vector<uint32_t> pos;
            for (uint32_t j = 0; j < ops[i].n; ++j)
                for (uint32_t k = 0; k < b.K; ++k)
                    if (b.qubits[k] == ops[i].qs[j]) pos.push_back(k);
            apply_to_block(b.m, b.K, pos, ops[i].m);
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
    std::vector<sqf::Op> ops;
    ops.reserve(ngates);
    for (uint32_t i = 0; i < ngates; ++i) {
        if (gates[i].nctrl > 3 || gates[i].target >= nq) return SUQA_ERR_ARG;
        if (gates[i].kind == SUQA_GATE_CX || gates[i].kind == SUQA_GATE_MCX)
            for (uint32_t k = 0; k < gates[i].nctrl; ++k)
                if (gates[i].ctrl[k] >= nq) return SUQA_ERR_ARG;
        ops.push_back(sqf::gate_op(gates[i]));
    }
    const auto blocks = sqf::fuse(ops, nq, width, 256);
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
