// tiled.hip -- multi-gate passes of the FP64 engine.
//
// With tile passes on (suqa_hip_set_tile_bits), the gate sweeps that the API
// calls produce (a pair operation or a phase rotation over a selection) are
// queued, and every operation that reads the state first flushes the queue.
// The flush cuts the queue, in program order, into passes: a pass is a run of
// up to kMaxPass consecutive gates whose partner masks X all lie inside one
// set T of t qubits (T always holds qubits 0..3, so a tile is a set of whole
// 128-byte lines).  Control and selection bits and Y/Z parity bits may lie
// anywhere: outside T they are constant over a tile and are tested once per
// tile.  A pass is one kernel: each workgroup loads a tile (the 2^t
// amplitudes that share the bits outside T) into LDS, applies the pass's
// gates one after the other with a barrier between gates, and writes the
// tile back, so the state crosses HBM once per pass instead of once per gate.
// Gates keep their order and every amplitude is computed with the one-gate
// sweeps' arithmetic (sweep_math.hpp), so results are bit-identical to
// applying the gates one sweep at a time.  A bit fixed to the same value by
// every gate of a pass (a common control) is also fixed in the tile
// enumeration: tiles outside the selection are never read.
#include <algorithm>
#include <cstring>

#include "engine.hpp"
#include "sweep_math.hpp"

namespace sqe {

namespace {

constexpr uint32_t kMaxPass = 24;
constexpr uint32_t kLowFixed = 4;   // qubits 0..3: 16 amplitudes = one 128-byte line

typedef double dv2t __attribute__((ext_vector_type(2)));

struct TileGate {
    uint32_t kind;        // OP_SWAP / OP_Y / OP_H / OP_PAULI or kPhaseOp
    uint32_t pivot;       // tile position of the lowest partner bit (pair operations)
    uint32_t xc, mc, vc;  // partner mask and in-tile selection, in tile positions
    uint32_t yzc;         // in-tile parity mask (Y|Z positions or the phase's zmask)
    uint32_t base_imag, base_neg;
    uint64_t mo, vo;      // selection bits outside the tile (global positions)
    uint64_t yzo;         // parity bits outside the tile
    double c, s;
};

struct TileArgs {
    uint64_t ntiles;
    uint64_t fixval;      // common selection values outside the tile
    uint32_t t;           // tile bits
    uint32_t nfix;        // positions fixed in the tile enumeration (tile bits + common selection)
    uint32_t ngates;
    uint32_t pad;
    uint8_t fixpos[64];   // ascending
    uint8_t tpos[16];     // tile bit positions, ascending (tile index order)
    TileGate g[kMaxPass];
};

// Global offset of tile element e (tile positions -> state bits).
__device__ __forceinline__ uint64_t tile_offset(uint32_t e, const TileArgs& a) {
    uint64_t o = 0;
    for (uint32_t j = 0; j < a.t; ++j)
        if ((e >> j) & 1u) o |= 1ull << a.tpos[j];
    return o;
}

template <int OP>
__device__ __forceinline__ void tile_pairs(double* lre, double* lim, uint32_t half, const TileGate& g, uint32_t opar) {
    PauliCoef pc;
    pc.c = g.c;
    pc.s = g.s;
    pc.yz = 0;
    pc.ny_odd = 0;
    pc.base_imag = g.base_imag;
    pc.base_neg = g.base_neg;
    const uint32_t lo = (1u << g.pivot) - 1u;
    for (uint32_t p = threadIdx.x; p < half; p += kBlock) {
        const uint32_t u = ((p & ~lo) << 1) | (p & lo);
        const uint32_t w = u ^ g.xc;
        uint32_t ia, ib;
        if ((u & g.mc) == g.vc) {
            ia = u;
            ib = w;
        } else if ((w & g.mc) == g.vc) {
            ia = w;
            ib = u;
        } else {
            continue;
        }
        const double ra = lre[ia], ma = lim[ia], rb = lre[ib], mb = lim[ib];
        // the partner of a is b and vice versa
        const uint32_t par_a = OP == OP_PAULI ? (((uint32_t)__popc(ib & g.yzc) & 1u) ^ opar) : 0u;
        const uint32_t par_b = OP == OP_PAULI ? (((uint32_t)__popc(ia & g.yzc) & 1u) ^ opar) : 0u;
        lre[ia] = pair_out_par<OP>(true, false, ra, rb, mb, par_a, pc);
        lim[ia] = pair_out_par<OP>(true, true, ma, mb, rb, par_a, pc);
        lre[ib] = pair_out_par<OP>(false, false, rb, ra, ma, par_b, pc);
        lim[ib] = pair_out_par<OP>(false, true, mb, ma, ra, par_b, pc);
    }
}

__global__ __launch_bounds__(kBlock) void k_tile(double* __restrict__ re, double* __restrict__ im, const TileArgs a) {
    extern __shared__ double lds[];
    const uint32_t tsz = 1u << a.t;
    double* lre = lds;
    double* lim = lds + tsz;
    const uint64_t G = gridDim.x, b = blockIdx.x;
    const uint64_t tile = (b & 7ull) * (G >> 3) + (b >> 3);   // XCD-contiguous (grid % 8 == 0)
    if (tile >= a.ntiles) return;                             // whole workgroup leaves together
    uint64_t base = tile;
    for (uint32_t j = 0; j < a.nfix; ++j) {
        const uint64_t lo = (1ull << a.fixpos[j]) - 1ull;
        base = (base & lo) | ((base & ~lo) << 1);
    }
    base |= a.fixval;

    // 1. tile -> LDS, 16-byte accesses, four in flight per array per thread
    constexpr uint32_t kU = 4;
    for (uint32_t e0 = 2 * threadIdx.x; e0 < tsz; e0 += 2 * kBlock * kU) {
        dv2t vr[kU], vi[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            const uint32_t e = e0 + u * 2 * kBlock;
            if (e < tsz) {
                const uint64_t g = base | tile_offset(e, a);
                vr[u] = __builtin_nontemporal_load(reinterpret_cast<const dv2t*>(re + g));
                vi[u] = __builtin_nontemporal_load(reinterpret_cast<const dv2t*>(im + g));
            }
        }
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            const uint32_t e = e0 + u * 2 * kBlock;
            if (e < tsz) {
                *reinterpret_cast<dv2t*>(lre + e) = vr[u];
                *reinterpret_cast<dv2t*>(lim + e) = vi[u];
            }
        }
    }
    __syncthreads();

    // 2. the gates, in order
    for (uint32_t k = 0; k < a.ngates; ++k) {
        const TileGate& g = a.g[k];
        if ((base & g.mo) != g.vo) continue;   // uniform: this tile is outside the gate's selection
        const uint32_t opar = (uint32_t)__popcll(base & g.yzo) & 1u;
        switch (g.kind) {
        case kPhaseOp:
            for (uint32_t e = threadIdx.x; e < tsz; e += kBlock) {
                if ((e & g.mc) != g.vc) continue;
                const uint32_t par = ((uint32_t)__popc(e & g.yzc) & 1u) ^ opar;
                const double s = par ? -g.s : g.s;
                const double r = lre[e], m = lim[e];
                lre[e] = phase_out(false, r, m, g.c, s);
                lim[e] = phase_out(true, m, r, g.c, s);
            }
            break;
        case OP_SWAP: tile_pairs<OP_SWAP>(lre, lim, tsz / 2, g, opar); break;
        case OP_Y: tile_pairs<OP_Y>(lre, lim, tsz / 2, g, opar); break;
        case OP_H: tile_pairs<OP_H>(lre, lim, tsz / 2, g, opar); break;
        default: tile_pairs<OP_PAULI>(lre, lim, tsz / 2, g, opar); break;
        }
        __syncthreads();
    }

    // 3. LDS -> tile
    for (uint32_t e = 2 * threadIdx.x; e < tsz; e += 2 * kBlock) {
        const uint64_t g = base | tile_offset(e, a);
        __builtin_nontemporal_store(*reinterpret_cast<const dv2t*>(lre + e), reinterpret_cast<dv2t*>(re + g));
        __builtin_nontemporal_store(*reinterpret_cast<const dv2t*>(lim + e), reinterpret_cast<dv2t*>(im + g));
    }
}

uint32_t compact(uint64_t m, const uint8_t* tpos, uint32_t t) {
    uint32_t o = 0;
    for (uint32_t j = 0; j < t; ++j)
        if ((m >> tpos[j]) & 1ull) o |= 1u << j;
    return o;
}

hipError_t run_single(suqa_hip_ctx* c, const GateDesc& d) {
    Sel sel;
    if (!make_sel(c->nq, d.M, d.V, &sel)) return hipSuccess;
    ++c->single_sweeps;
    if (d.kind == kPhaseOp) return launch_phase_now(c, sel, d.c, d.s, d.zmask, d.tag, d.alg_bytes);
    return launch_pair_now(c, d.kind, sel, d.X, d.pc, d.tag, d.alg_bytes);
}

hipError_t run_pass(suqa_hip_ctx* c, const GateDesc* q, uint32_t n, uint64_t T, uint32_t t) {
    const uint32_t nq = c->nq;
    for (uint32_t b = 0; b < nq && (uint32_t)__builtin_popcountll(T) < t; ++b) T |= 1ull << b;   // pad
    TileArgs a;
    std::memset(&a, 0, sizeof(a));
    a.t = t;
    uint32_t k = 0;
    for (uint32_t b = 0; b < nq; ++b)
        if ((T >> b) & 1ull) a.tpos[k++] = (uint8_t)b;
    a.ngates = n;
    uint64_t mcom = ~0ull, vcom = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const GateDesc& d = q[i];
        TileGate& g = a.g[i];
        g.kind = (uint32_t)d.kind;
        g.mc = compact(d.M, a.tpos, t);
        g.vc = compact(d.V, a.tpos, t);
        g.mo = d.M & ~T;
        g.vo = d.V & ~T;
        const uint64_t par = d.kind == kPhaseOp ? d.zmask : d.pc.yz;
        g.yzc = compact(par, a.tpos, t);
        g.yzo = par & ~T;
        if (d.kind == kPhaseOp) {
            g.c = d.c;
            g.s = d.s;
        } else {
            g.xc = compact(d.X, a.tpos, t);
            g.pivot = (uint32_t)__builtin_ctz(g.xc);
            g.c = d.pc.c;
            g.s = d.pc.s;
            g.base_imag = d.pc.base_imag;
            g.base_neg = d.pc.base_neg;
        }
        // bits outside the tile that every gate fixes to the same value
        if (i == 0) vcom = g.vo;
        mcom &= g.mo;
        mcom &= ~((g.vo ^ vcom) & mcom);
    }
    vcom &= mcom;
    const uint64_t fix = T | mcom;
    a.nfix = 0;
    for (uint32_t b = 0; b < nq; ++b)
        if ((fix >> b) & 1ull) a.fixpos[a.nfix++] = (uint8_t)b;
    a.fixval = vcom;
    a.ntiles = 1ull << (nq - a.nfix);
    const uint64_t g8 = ((a.ntiles + 7) / 8) * 8;
    const size_t lds = sizeof(double) * 2 * ((size_t)1 << t);
    ++c->passes;
    c->pass_gates += n;
    timing_begin(c, "tile_pass", 32.0 * (double)((uint64_t)1 << t) * (double)a.ntiles);
    hipLaunchKernelGGL(k_tile, dim3((uint32_t)g8), dim3(kBlock), lds, c->stream, c->d_re, c->d_im, a);
    timing_end(c);
    return hipGetLastError();
}

}  // namespace

hipError_t flush(suqa_hip_ctx* c) {
    if (c->queue.empty()) return hipSuccess;
    std::vector<GateDesc> q;
    q.swap(c->queue);
    const uint32_t t = std::min(c->tile_bits, c->nq);
    const uint32_t f = std::min(kLowFixed, t);
    const uint64_t low = (1ull << f) - 1ull;
    size_t i = 0;
    while (i < q.size()) {
        uint64_t T = low;
        size_t j = i;
        while (j < q.size() && j - i < kMaxPass) {
            const uint64_t need = q[j].kind == kPhaseOp ? 0ull : q[j].X;
            if ((uint32_t)__builtin_popcountll(T | need) > t) break;
            T |= need;
            ++j;
        }
        hipError_t e;
        if (j - i <= 1) {   // a lone gate (or one whose partner bits exceed the tile): one sweep
            e = run_single(c, q[i]);
            j = i + 1;
        } else {
            e = run_pass(c, q.data() + i, (uint32_t)(j - i), T, t);
        }
        if (e != hipSuccess) return e;
        i = j;
    }
    return hipSuccess;
}

}  // namespace sqe
