// tiled.hip -- multi-gate passes of the FP64 engine (and of the FP32
// engine's gate-pass mode, on a float state with the same kernels).
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
// gates in order, and writes the tile back, so the state crosses HBM once per
// pass instead of once per gate.
//
// Inside a pass, consecutive gates whose flipped bits fit in 3 tile bits F
// form a register step: each thread takes the 8 amplitudes that F spans
// (a group) from LDS into registers, applies the step's gates to them one
// after the other and writes them back, so LDS is read and written once per
// step.  A gate's selection and parity split into a part on F, the same for
// every group and precomputed on the host as 8-bit tables over the group's 8
// elements, and a part on the other tile bits, tested once per group; the
// per-element work is the gate arithmetic alone.  A gate that flips more than
// 3 bits gets a step of its own, applied pair by pair in LDS.  A state that
// is one tile of at most 2^10 amplitudes (a latency-bound workgroup) runs
// k_tile_tiny instead: one amplitude per thread, one exchange per gate.
//
// Gates keep their order and every amplitude is computed with the one-gate
// sweeps' arithmetic (sweep_math.hpp), so results are bit-identical to
// applying the gates one sweep at a time.  A bit fixed to the same value by
// every gate of a pass (a common control) is also fixed in the tile
// enumeration: tiles outside the selection are never read.
#include <algorithm>
#include <cstring>
#include <type_traits>
#include <utility>

#include "deposit.hpp"
#include "engine.hpp"
#include "lanes.hpp"
#include "sweep_math.hpp"

namespace sqe {

namespace {

constexpr uint32_t kMaxPass = kPassMaxGates;
constexpr uint32_t kLowFixed = kPassLowFixed;     // qubits 0..3: 16 amplitudes, one 128-byte line in FP64
constexpr uint32_t kSub = 3, kLoc = 1u << kSub;   // register step: 3 bits, 8 amplitudes per thread
constexpr uint32_t kTinyBits = 10;                // one-tile states up to 2^10: k_tile_tiny
constexpr uint32_t kMaxRuns = 20;                 // runs of consecutive set bits of a mask of <= 39 bits

// Amplitude types: double (the FP64 engine) or float (the FP32 engine's gate
// passes).  A tile moves between HBM and LDS in 16-byte chunks of kVec
// amplitudes per array; tiles have at most kMaxBits bits.
template <typename T>
struct AmpT;
template <>
struct AmpT<double> {
    typedef double v8 __attribute__((ext_vector_type(8)));   // a thread's 8 register amplitudes
    typedef double chunk __attribute__((ext_vector_type(2)));
    static constexpr uint32_t kVec = 2, kMaxBits = kMaxTileBits;
};
template <>
struct AmpT<float> {
    typedef float v8 __attribute__((ext_vector_type(8)));
    typedef float chunk __attribute__((ext_vector_type(4)));
    static constexpr uint32_t kVec = 4, kMaxBits = kMaxTileBits32;
};
// A tile in LDS: element e is the amplitude (re, im) of tile index e, one
// 8-byte (FP32) or 16-byte (FP64) access.
template <typename T>
using Cplx = T __attribute__((ext_vector_type(2)));

// What a register step reads of a gate, first in TileGate so that the gate
// loop fetches it with one 64-byte scalar load (plus 16 bytes for FP64
// coefficients) and one wait, instead of field by field.
struct GateR {
    uint64_t mo, vo;      // selection bits outside the tile (global positions)
    uint64_t yzo;         // parity bits outside the tile
    // the part on the tile bits off the step's bits F (tile positions, zero on F) ...
    uint32_t mcg, vcg, yzg;
    uint32_t arm;         // register-step code path (reg_gate)
    // ... and the part on F
    uint32_t lsel;        // bit j: group element j meets the selection's part on F
    uint32_t lpar;        // bit j: parity of element j's bits on F in the parity mask
    uint32_t gtriv;       // mcg == 0: every group meets the selection
    uint32_t base_neg;
    float cf, sf;         // c, s rounded to FP32 (float passes)
    double c, s;
};
static_assert(sizeof(GateR) == 80, "GateR is read as 64 + 16 bytes");

// The LDS-step fields in 16 and 8 bits (tile positions are below 2^13), so
// that kMaxPass descriptors fit the 4 KiB of kernel arguments.
struct TileGate : GateR {
    uint16_t xc, mc, vc;  // partner mask and selection in tile positions (LDS steps)
    uint16_t yzc;         // parity mask in tile positions (Y|Z positions or the phase's zmask)
    uint8_t kind;         // OP_SWAP / OP_Y / OP_H / OP_PAULI or kPhaseOp
    uint8_t pivot;        // tile position of the lowest partner bit (pair operations)
    uint8_t xl;           // partner mask over F
    uint8_t dense;        // gtriv, and every element (pair) of a group is selected
    uint8_t base_imag;
};
static_assert(sizeof(TileGate) == 96, "TileGate layout");

struct TileArgs {
    uint64_t ntiles;
    uint64_t fixval;      // common selection values outside the tile
    uint32_t t;           // tile bits
    uint32_t ngates;
    uint32_t nsteps;
    uint64_t tmask;       // the tile's state bits (tile element e -> deposit(e, tmask))
    uint64_t freemask;    // state bits enumerated by the tile number (not in the tile,
                          // not fixed by a selection common to every gate)
    // The same two maps as runs of consecutive mask bits, for deposit_runs (the
    // bit-by-bit deposit costs ~10 scalar instructions per mask bit, a few
    // hundred per tile): run i = lo | len << 8 | src << 16.
    uint32_t nfree_runs, ntile_runs;
    uint32_t free_runs[kMaxRuns];
    uint32_t tile_runs[kMaxRuns];   // only the runs fed by the bits of a thread's first element
    uint64_t kbit[3];     // state offsets of the tile bits that chunk index bits 0, 1, 2 stand for
    // span tiles: tile element e sits at base ^ (XOR of lvec[m] over e's set
    // bits m); skew = 0 when every lvec is one qubit (then tmask and the runs
    // give the same offsets)
    uint32_t skew;
    uint64_t lvec[kMaxTileBits32];
    struct Step {
        uint32_t f[kSub]; // tile positions of the step's bits, ascending; f[0] = kLdsStep: LDS step
        uint32_t g0, g1;  // gates [g0, g1)
    } step[kMaxPass];
    TileGate g[kMaxPass];
};
constexpr uint32_t kLdsStep = 0xffu;
static_assert(sizeof(TileArgs) <= 4096, "kernel arguments are limited to 4 KiB");

// A gate's coefficients in the pass's amplitude type (FP64: the exact values).
template <typename T>
__device__ __forceinline__ T gate_c(const GateR& g);
template <>
__device__ __forceinline__ double gate_c<double>(const GateR& g) { return g.c; }
template <>
__device__ __forceinline__ float gate_c<float>(const GateR& g) { return g.cf; }
template <typename T>
__device__ __forceinline__ T gate_s(const GateR& g);
template <>
__device__ __forceinline__ double gate_s<double>(const GateR& g) { return g.s; }
template <>
__device__ __forceinline__ float gate_s<float>(const GateR& g) { return g.sf; }

template <typename T>
struct Coef {   // what pair_out_par reads of a pair gate
    T c, s;
    uint32_t base_imag, base_neg;
};
template <typename T>
__device__ __forceinline__ Coef<T> coef(const TileGate& g) {
    return Coef<T>{gate_c<T>(g), gate_s<T>(g), g.base_imag, g.base_neg};
}

// LDS step: pair gate over the whole tile, pair (u, u ^ X) per thread.
template <int OP, typename T>
__device__ __forceinline__ void tile_pairs(Cplx<T>* lc, uint32_t half, const TileGate& g, uint32_t opar) {
    const Coef<T> pc = coef<T>(g);
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
        const Cplx<T> va = lc[ia], vb = lc[ib];
        const T ra = va.x, ma = va.y, rb = vb.x, mb = vb.y;
        // the partner of a is b and vice versa
        const uint32_t par_a = OP == OP_PAULI ? (((uint32_t)__popc(ib & g.yzc) & 1u) ^ opar) : 0u;
        const uint32_t par_b = OP == OP_PAULI ? (((uint32_t)__popc(ia & g.yzc) & 1u) ^ opar) : 0u;
        Cplx<T> oa, ob;
        oa.x = pair_out_par<OP>(true, false, ra, rb, mb, par_a, pc);
        oa.y = pair_out_par<OP>(true, true, ma, mb, rb, par_a, pc);
        ob.x = pair_out_par<OP>(false, false, rb, ra, ma, par_b, pc);
        ob.y = pair_out_par<OP>(false, true, mb, ma, ra, par_b, pc);
        lc[ia] = oa;
        lc[ib] = ob;
    }
}

// A thread's 8 register amplitudes.  Gate steps take and return them by
// value: with references, the compiler merges the stores of the switch cases
// below into one store through a pointer phi and the amplitudes end up in
// scratch memory.
template <typename T>
struct RM {
    Cplx<T> e[kLoc];   // (re, im) pairs: one register pair per amplitude, as LDS holds them
};

// Register steps compute every element and pick results with selects, never
// with branches on data: a branch around an update makes the compiler copy
// the whole register set at the join.
//
// Every arm updates the 8 amplitudes in place: each output is defined after
// the last use of the input it replaces (products and sums of all inputs
// first, then the final operation of each output, or a select whose other
// operand is that input).  The dispatch's join then needs no copy: the
// results are already in the registers the next gate reads.  One arm that
// computes an output while its input is still live makes the compiler keep
// two register sets and copy all 16 amplitudes after every gate.

// A copy of v in a new register (passed-through values in the select arms).
__device__ __forceinline__ float fresh(float v) {
    float r;
    asm("v_mov_b32 %0, %1" : "=v"(r) : "v"(v));
    return r;
}
__device__ __forceinline__ double fresh(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    uint32_t lo = (uint32_t)b, hi = (uint32_t)(b >> 32), rl, rh;
    asm("v_mov_b32 %0, %1" : "=v"(rl) : "v"(lo));
    asm("v_mov_b32 %0, %1" : "=v"(rh) : "v"(hi));
    return __longlong_as_double((long long)(((uint64_t)rh << 32) | rl));
}
// Makes the values exist before anything that uses them is computed (no
// instruction: an empty asm reading and writing them).  The arms compute
// every partial result from the inputs first and pass them through here, so
// the outputs, computed from the barrier's results, are defined after the
// inputs' last uses whatever order the scheduler picks.
template <typename T>
__device__ __forceinline__ void settle(T& a, T& b, T& c, T& d) {
    asm volatile("" : "+v"(a), "+v"(b), "+v"(c), "+v"(d));
}

template <typename T>
__device__ __forceinline__ void settle6(T& a, T& b, T& c, T& d, T& e, T& f) {
    asm volatile("" : "+v"(a), "+v"(b), "+v"(c), "+v"(d), "+v"(e), "+v"(f));
}

// Keeps the addends of the FP32 FMAs alive past them: an FMA whose addend
// dies is emitted in the accumulating form (result in the addend's register),
// which moves the amplitude out of its own register.
template <typename T>
__device__ __forceinline__ void after(T a, T b, T c, T d) {
    asm volatile("" ::"v"(a), "v"(b), "v"(c), "v"(d));
}

// Exchanges a and b in their registers.
__device__ __forceinline__ void vswap(float& a, float& b) { asm("v_swap_b32 %0, %1" : "+v"(a), "+v"(b)); }
__device__ __forceinline__ void vswap(double& a, double& b) {
    const uint64_t ua = (uint64_t)__double_as_longlong(a), ub = (uint64_t)__double_as_longlong(b);
    uint32_t al = (uint32_t)ua, ah = (uint32_t)(ua >> 32), bl = (uint32_t)ub, bh = (uint32_t)(ub >> 32);
    asm("v_swap_b32 %0, %1" : "+v"(al), "+v"(bl));
    asm("v_swap_b32 %0, %1" : "+v"(ah), "+v"(bh));
    a = __longlong_as_double((long long)(((uint64_t)ah << 32) | al));
    b = __longlong_as_double((long long)(((uint64_t)bh << 32) | bl));
}

// FP32 arms fuse each product and sum into one FMA (more accurate, fewer
// instructions); FP64 arms keep the reference's separate rounding of each
// operation, so FP64 results stay bit-identical.
template <typename T>
constexpr bool kFma = std::is_same<T, float>::value;

// Pair operations of the register arms (OPV): the four pair kinds, with the
// Pauli rotation split by its base (real: psi' = c psi + t psi_partner,
// imaginary: the partner's other component, re with -t, im with +t).
constexpr int kSwap = 0, kY = 1, kH = 2, kPauliR = 3, kPauliI = 4;
// Selection modes: every pair (kDense), the pairs whose element u has tile
// bit CB equal to CV (kHalf + 2 CB + CV, no other selection), or per element
// from lsel and gpass (kGeneral).
constexpr int kDense = 0, kGeneral = 1, kHalf = 2;
// kGroup: every element of the groups that meet the selection's bits off F
// (and none on F): the dense update on the lanes of those groups only.
constexpr int kGroup = 8;
// Added to a Pauli or phase mode: the gate has no parity qubits at all (no Y or
// Z in its string, no zmask): its sign is uniform and no element tests lpar.
constexpr int kNoPar = 16;

// Pauli sign: t = (base_neg ^ parity) ? -s : s; tg, ntg are the signed s and
// its negation for a parity-0 element of this group, flipped where the
// element's own bits on F have odd parity (lpar).
template <typename T>
__device__ __forceinline__ T pauli_t(const GateR& g, int elem, T tg, T ntg) {
    return ((g.lpar >> elem) & 1u) ? ntg : tg;
}

// Pair (u, u ^ XL) of a register step, u = J when J lacks the lowest bit of
// XL.  Dense pairs are updated in place; general pairs compute their outputs
// first and select them per element.  Outputs follow pair_out_par exactly
// (same operations, same order), so FP64 results stay bit-identical.
template <int OPV, int XL, int SEL, int J, typename T>
__device__ __forceinline__ RM<T> reg_pair_one(RM<T> x, const GateR& g, T c, T tg, T ntg, bool act) {
    constexpr int PIV = XL & -XL;
    constexpr int SM = SEL & (kNoPar - 1);
    constexpr bool NP = (SEL & kNoPar) != 0;
    if constexpr ((J & PIV) == 0) {
        constexpr int u = J, w = J ^ XL;
        constexpr bool dense = SM == kDense || (SM >= kHalf && (((u >> ((SM - kHalf) >> 1)) & 1) == ((SM - kHalf) & 1)));
        if constexpr (SM >= kHalf && !dense) return x;   // this pair is outside the half selection
        T ru = x.e[u].x, mu = x.e[u].y, rw = x.e[w].x, mw = x.e[w].y;
        constexpr T kh = (T)kTwoSqInv;
        if constexpr (dense) {
            if constexpr (OPV == kSwap) {
                vswap(ru, rw);
                vswap(mu, mw);
            } else if constexpr (OPV == kY) {   // re_u' = im_w, im_w' = re_u, im_u' = -re_w, re_w' = -im_u
                vswap(ru, mw);
                vswap(mu, rw);
                mu = mu * (T)-1.0;
                rw = rw * (T)-1.0;
            } else if constexpr (OPV == kH) {   // u: k (u + w); w: k (u - w)
                T sr = ru + rw, sm = mu + mw, dr = ru - rw, dm = mu - mw;
                settle(sr, sm, dr, dm);
                ru = kh * sr;
                mu = kh * sm;
                rw = kh * dr;
                mw = kh * dm;
            } else {
                const T tu = NP ? tg : pauli_t(g, w, tg, ntg), tw = NP ? tg : pauli_t(g, u, tg, ntg);
                T pru = (T)0, pmu = (T)0, prw = (T)0, pmw = (T)0;
                if constexpr (!kFma<T>) {
                    pru = ru * c;
                    pmu = mu * c;
                    prw = rw * c;
                    pmw = mw * c;
                }
                T qru, qmu, qrw, qmw;
                if constexpr (OPV == kPauliR) {
                    qru = rw * tu;
                    qmu = mw * tu;
                    qrw = ru * tw;
                    qmw = mu * tw;
                } else {
                    qru = mw * (-tu);
                    qmu = rw * tu;
                    qrw = mu * (-tw);
                    qmw = ru * tw;
                }
                if constexpr (kFma<T>) {   // FP32: c psi + q in one rounding
                    settle(qru, qmu, qrw, qmw);
                    ru = __builtin_fmaf(ru, c, qru);
                    mu = __builtin_fmaf(mu, c, qmu);
                    rw = __builtin_fmaf(rw, c, qrw);
                    mw = __builtin_fmaf(mw, c, qmw);
                    after(qru, qmu, qrw, qmw);
                } else {
                    settle(pru, pmu, prw, pmw);
                    settle(qru, qmu, qrw, qmw);
                    ru = pru + qru;
                    mu = pmu + qmu;
                    rw = prw + qrw;
                    mw = pmw + qmw;
                }
            }
        } else {   // kGeneral
            // no short-circuit: a branch here would split the arm into blocks
            bool on;
            if constexpr (XL == PIV) {
                on = (((g.lsel >> u) & 1u) != 0u) & act;
            } else {   // several flipped bits (SWAP, Pauli strings): either element may be selected
                on = ((((g.lsel >> u) | (g.lsel >> w)) & 1u) != 0u) & act;
            }
            T cru, cmu, crw, cmw;
            if constexpr (OPV == kSwap) {
                const T tr = fresh(ru), tm = fresh(mu);
                ru = on ? rw : tr;
                rw = on ? tr : rw;
                mu = on ? mw : tm;
                mw = on ? tm : mw;
            } else {
                if constexpr (OPV == kY) {
                    cru = fresh(mw);
                    cmw = fresh(ru);
                    cmu = rw * (T)-1.0;
                    crw = mu * (T)-1.0;
                } else if constexpr (OPV == kH) {
                    cru = kh * (ru + rw);
                    cmu = kh * (mu + mw);
                    crw = kh * (ru - rw);
                    cmw = kh * (mu - mw);
                } else {
                    const T tu = NP ? tg : pauli_t(g, w, tg, ntg), tw = NP ? tg : pauli_t(g, u, tg, ntg);
                    if constexpr (kFma<T>) {
                        if constexpr (OPV == kPauliR) {
                            cru = __builtin_fmaf(ru, c, rw * tu);
                            cmu = __builtin_fmaf(mu, c, mw * tu);
                            crw = __builtin_fmaf(rw, c, ru * tw);
                            cmw = __builtin_fmaf(mw, c, mu * tw);
                        } else {
                            cru = __builtin_fmaf(ru, c, mw * (-tu));
                            cmu = __builtin_fmaf(mu, c, rw * tu);
                            crw = __builtin_fmaf(rw, c, mu * (-tw));
                            cmw = __builtin_fmaf(mw, c, ru * tw);
                        }
                    } else if constexpr (OPV == kPauliR) {
                        cru = ru * c + rw * tu;
                        cmu = mu * c + mw * tu;
                        crw = rw * c + ru * tw;
                        cmw = mw * c + mu * tw;
                    } else {
                        cru = ru * c + mw * (-tu);
                        cmu = mu * c + rw * tu;
                        crw = rw * c + mu * (-tw);
                        cmw = mw * c + ru * tw;
                    }
                }
                settle(cru, cmu, crw, cmw);
                ru = on ? cru : ru;
                mu = on ? cmu : mu;
                rw = on ? crw : rw;
                mw = on ? cmw : mw;
            }
        }
        x.e[u].x = ru;
        x.e[u].y = mu;
        x.e[w].x = rw;
        x.e[w].y = mw;
    }
    return x;
}

template <int OPV, int XL, int SEL, typename T, int... J>
__device__ __forceinline__ RM<T> reg_pairs_seq(RM<T> x, const GateR& g, bool gpass, uint32_t gpar,
                                               std::integer_sequence<int, J...>) {
    constexpr bool NP = (SEL & kNoPar) != 0;
    T c = (T)0, tg = (T)0, ntg = (T)0;
    if constexpr (OPV == kPauliR || OPV == kPauliI) {
        c = gate_c<T>(g);
        const T s = gate_s<T>(g);
        tg = ((NP ? 0u : gpar) ^ g.base_neg) ? -s : s;
        ntg = -tg;
    }
    if constexpr ((SEL & (kNoPar - 1)) == kGroup) {
        if (gpass) {   // per lane: the update runs under the lanes' mask, in place
            ((x = reg_pair_one<OPV, XL, kDense | (SEL & kNoPar), J>(x, g, c, tg, ntg, true)), ...);
        }
        return x;
    } else {
        const bool act = (g.gtriv != 0u) | gpass;   // this group meets the selection's bits off F
        ((x = reg_pair_one<OPV, XL, SEL, J>(x, g, c, tg, ntg, act)), ...);
        return x;
    }
}

// Phase rotation of the register amplitudes (util_rotate2 on each selected
// one): dense (all 8), half (the elements whose bit B equals V), general.
template <int SEL, int J, typename T>
__device__ __forceinline__ RM<T> reg_phase_one(RM<T> x, const GateR& g, T c, T sg, T nsg, bool act) {
    constexpr int SM = SEL & (kNoPar - 1);
    constexpr bool NP = (SEL & kNoPar) != 0;
    constexpr bool half_on = SM >= kHalf && (((J >> ((SM - kHalf) >> 1)) & 1) == ((SM - kHalf) & 1));
    if constexpr (SM >= kHalf && !half_on) return x;
    const T sn = NP ? sg : (((g.lpar >> J) & 1u) ? nsg : sg);
    const T rr = x.e[J].x, mm = x.e[J].y;
    // the next element's inputs pass the same barrier: its products start
    // after this element's, which bounds the registers in flight
    constexpr int N = (J + 1) & 7;
    T nr0 = x.e[N].x, nm0 = x.e[N].y;
    T nr, nm;
    if constexpr (kFma<T>) {   // FP32: re c - im s and im c + re s, one rounding each
        T b = mm * sn, d = rr * sn;
        settle(b, d, nr0, nm0);
        nr = __builtin_fmaf(rr, c, -b);
        nm = __builtin_fmaf(mm, c, d);
        after(b, d, b, d);
    } else {
        T a = rr * c, b = mm * sn, cc = mm * c, d = rr * sn;
        settle6(a, b, cc, d, nr0, nm0);
        nr = a - b;
        nm = cc + d;
    }
    x.e[N].x = nr0;
    x.e[N].y = nm0;
    if constexpr (SM == kGeneral) {
        const bool on = (((g.lsel >> J) & 1u) != 0u) & act;
        x.e[J].x = on ? nr : rr;
        x.e[J].y = on ? nm : mm;
    } else {
        x.e[J].x = nr;
        x.e[J].y = nm;
    }
    return x;
}

template <int SEL, typename T, int... J>
__device__ __forceinline__ RM<T> reg_phase_seq(RM<T> x, const GateR& g, bool gpass, uint32_t gpar,
                                               std::integer_sequence<int, J...>) {
    constexpr bool NP = (SEL & kNoPar) != 0;
    const T c = gate_c<T>(g), s = gate_s<T>(g);
    const T sg = (NP ? 0u : gpar) ? -s : s, nsg = -sg;
    if constexpr ((SEL & (kNoPar - 1)) == kGroup) {
        if (gpass) {
            ((x = reg_phase_one<kDense | (SEL & kNoPar), J>(x, g, c, sg, nsg, true)), ...);
        }
        return x;
    } else {
        const bool act = (g.gtriv != 0u) | gpass;
        ((x = reg_phase_one<SEL, J>(x, g, c, sg, nsg, act)), ...);
        return x;
    }
}

// One flat dispatch over every register-step code path, on the arm index the
// host computes (local_tables): nested switches with a fall-through from the
// dense cases made the compiler copy the 16 amplitudes into the result
// registers before dispatching, on top of the copy back at the join.
// Every arm is its own case: an earlier dense dispatch written as `case 1`,
// `case 2`, `default` (= 4) was miscompiled by hipcc (ROCm 7.2), which merged
// the OP_Y cases and left half of the xl = 4 outputs unassigned.
// Arms (kArm* below): phase dense / general / half (bit b, value v); per pair
// kind: dense one-bit partner masks 1, 2, 4; general one-bit masks; general
// multi-bit masks 3, 5, 6, 7 (SWAP and Pauli only); for SWAP also half arms
// (one-bit mask, control bit cb of F at value cv).
constexpr uint32_t kArmPhaseDense = 0, kArmPhaseGeneral = 1, kArmPhaseHalf = 2;   // + 2 b + v
constexpr uint32_t kArmPair = 8;              // + 16 OPV + ...
constexpr uint32_t kArmPairDense = 0;         // + log2 xl
constexpr uint32_t kArmPairGeneral1 = 3;      // + log2 xl
constexpr uint32_t kArmPairMulti = 6;         // + {3: 0, 5: 1, 6: 2, 7: 3}
constexpr uint32_t kArmSwapHalf = kArmPair + 16 * 5;   // + 4 log2 xl + 2 (cb index among the other two bits) + cv
constexpr uint32_t kArmPhaseGroup = kArmSwapHalf + 12;
constexpr uint32_t kArmPairGroup = kArmPhaseGroup + 1;   // + 3 OPV + log2 xl
// gates with no parity qubits: phase dense, group, half (+ 2 b + v); Pauli
// dense and group (+ 3 (imaginary base) + log2 xl)
constexpr uint32_t kArmPhaseNp = kArmPairGroup + 15;      // dense; + 1 group; + 2 + 2 b + v half
constexpr uint32_t kArmPauliNpDense = kArmPhaseNp + 8;
constexpr uint32_t kArmPauliNpGroup = kArmPauliNpDense + 6;

// G groups of a thread go through one dispatch: the gate's uniform work (its
// descriptor loads, the dispatch, the per-element selection and parity masks)
// is done once for all of them.
template <typename T, int G>
struct RMG {
    RM<T> q[G];
};
template <int G>
struct GroupSel {
    bool pass[G];      // the group meets the selection's bits off F
    uint32_t par[G];   // parity of the group's bits off F (and outside the tile)
};
#define SQE_EACH(CALL)                                                                                       \
    {                                                                                                        \
        _Pragma("unroll") for (int q_ = 0; q_ < G; ++q_) x.q[q_] = CALL;                                     \
        return x;                                                                                            \
    }
#define SQE_PAIR_ARMS(OPV)                                                                                   \
    case kArmPair + 16 * OPV + 0: SQE_EACH((reg_pairs_seq<OPV, 1, kDense>(x.q[q_], g, p.pass[q_], p.par[q_], seq)))   \
    case kArmPair + 16 * OPV + 1: SQE_EACH((reg_pairs_seq<OPV, 2, kDense>(x.q[q_], g, p.pass[q_], p.par[q_], seq)))   \
    case kArmPair + 16 * OPV + 2: SQE_EACH((reg_pairs_seq<OPV, 4, kDense>(x.q[q_], g, p.pass[q_], p.par[q_], seq)))   \
    case kArmPair + 16 * OPV + 3: SQE_EACH((reg_pairs_seq<OPV, 1, kGeneral>(x.q[q_], g, p.pass[q_], p.par[q_], seq))) \
    case kArmPair + 16 * OPV + 4: SQE_EACH((reg_pairs_seq<OPV, 2, kGeneral>(x.q[q_], g, p.pass[q_], p.par[q_], seq))) \
    case kArmPair + 16 * OPV + 5: SQE_EACH((reg_pairs_seq<OPV, 4, kGeneral>(x.q[q_], g, p.pass[q_], p.par[q_], seq)))
#define SQE_MULTI_ARMS(OPV)                                                                                  \
    case kArmPair + 16 * OPV + 6: SQE_EACH((reg_pairs_seq<OPV, 3, kGeneral>(x.q[q_], g, p.pass[q_], p.par[q_], seq))) \
    case kArmPair + 16 * OPV + 7: SQE_EACH((reg_pairs_seq<OPV, 5, kGeneral>(x.q[q_], g, p.pass[q_], p.par[q_], seq))) \
    case kArmPair + 16 * OPV + 8: SQE_EACH((reg_pairs_seq<OPV, 6, kGeneral>(x.q[q_], g, p.pass[q_], p.par[q_], seq))) \
    case kArmPair + 16 * OPV + 9: SQE_EACH((reg_pairs_seq<OPV, 7, kGeneral>(x.q[q_], g, p.pass[q_], p.par[q_], seq)))
// SWAP half arms: partner mask XL, control bit CB (the other two F bits) at CV
#define SQE_SWAP_HALF(LXL, XL, K, CB)                                                                        \
    case kArmSwapHalf + 4 * LXL + 2 * K + 0:                                                                 \
        SQE_EACH((reg_pairs_seq<kSwap, XL, kHalf + 2 * CB + 0>(x.q[q_], g, p.pass[q_], p.par[q_], seq)))     \
    case kArmSwapHalf + 4 * LXL + 2 * K + 1:                                                                 \
        SQE_EACH((reg_pairs_seq<kSwap, XL, kHalf + 2 * CB + 1>(x.q[q_], g, p.pass[q_], p.par[q_], seq)))
#define SQE_PHASE_ARM(CASE, SEL) case CASE: SQE_EACH((reg_phase_seq<SEL>(x.q[q_], g, p.pass[q_], p.par[q_], seq)))
#define SQE_NP_ARMS(OPV, I)                                                                                  \
    case kArmPauliNpDense + 3 * I + 0: SQE_EACH((reg_pairs_seq<OPV, 1, kDense | kNoPar>(x.q[q_], g, p.pass[q_], p.par[q_], seq)))  \
    case kArmPauliNpDense + 3 * I + 1: SQE_EACH((reg_pairs_seq<OPV, 2, kDense | kNoPar>(x.q[q_], g, p.pass[q_], p.par[q_], seq)))  \
    case kArmPauliNpDense + 3 * I + 2: SQE_EACH((reg_pairs_seq<OPV, 4, kDense | kNoPar>(x.q[q_], g, p.pass[q_], p.par[q_], seq)))  \
    case kArmPauliNpGroup + 3 * I + 0: SQE_EACH((reg_pairs_seq<OPV, 1, kGroup | kNoPar>(x.q[q_], g, p.pass[q_], p.par[q_], seq)))  \
    case kArmPauliNpGroup + 3 * I + 1: SQE_EACH((reg_pairs_seq<OPV, 2, kGroup | kNoPar>(x.q[q_], g, p.pass[q_], p.par[q_], seq)))  \
    case kArmPauliNpGroup + 3 * I + 2: SQE_EACH((reg_pairs_seq<OPV, 4, kGroup | kNoPar>(x.q[q_], g, p.pass[q_], p.par[q_], seq)))
#define SQE_GROUP_ARMS(OPV)                                                                                  \
    case kArmPairGroup + 3 * OPV + 0: SQE_EACH((reg_pairs_seq<OPV, 1, kGroup>(x.q[q_], g, p.pass[q_], p.par[q_], seq)))   \
    case kArmPairGroup + 3 * OPV + 1: SQE_EACH((reg_pairs_seq<OPV, 2, kGroup>(x.q[q_], g, p.pass[q_], p.par[q_], seq)))   \
    case kArmPairGroup + 3 * OPV + 2: SQE_EACH((reg_pairs_seq<OPV, 4, kGroup>(x.q[q_], g, p.pass[q_], p.par[q_], seq)))

template <typename T, int G>
__device__ __forceinline__ RMG<T, G> reg_gate(RMG<T, G> x, const GateR& g, const GroupSel<G>& p) {
    constexpr auto seq = std::make_integer_sequence<int, (int)kLoc>{};
    switch (g.arm) {
    SQE_PHASE_ARM(kArmPhaseDense, kDense)
    SQE_PHASE_ARM(kArmPhaseGeneral, kGeneral)
    SQE_PHASE_ARM(kArmPhaseHalf + 0, kHalf + 0)
    SQE_PHASE_ARM(kArmPhaseHalf + 1, kHalf + 1)
    SQE_PHASE_ARM(kArmPhaseHalf + 2, kHalf + 2)
    SQE_PHASE_ARM(kArmPhaseHalf + 3, kHalf + 3)
    SQE_PHASE_ARM(kArmPhaseHalf + 4, kHalf + 4)
    SQE_PHASE_ARM(kArmPhaseHalf + 5, kHalf + 5)
    SQE_PAIR_ARMS(kSwap)
    SQE_MULTI_ARMS(kSwap)
    SQE_PAIR_ARMS(kY)
    SQE_PAIR_ARMS(kH)
    SQE_PAIR_ARMS(kPauliR)
    SQE_MULTI_ARMS(kPauliR)
    SQE_PAIR_ARMS(kPauliI)
    SQE_MULTI_ARMS(kPauliI)
    SQE_SWAP_HALF(0, 1, 0, 1)
    SQE_SWAP_HALF(0, 1, 1, 2)
    SQE_SWAP_HALF(1, 2, 0, 0)
    SQE_SWAP_HALF(1, 2, 1, 2)
    SQE_SWAP_HALF(2, 4, 0, 0)
    SQE_SWAP_HALF(2, 4, 1, 1)
    SQE_PHASE_ARM(kArmPhaseGroup, kGroup)
    SQE_GROUP_ARMS(kSwap)
    SQE_GROUP_ARMS(kY)
    SQE_GROUP_ARMS(kH)
    SQE_GROUP_ARMS(kPauliR)
    SQE_GROUP_ARMS(kPauliI)
    SQE_PHASE_ARM(kArmPhaseNp + 0, kDense | kNoPar)
    SQE_PHASE_ARM(kArmPhaseNp + 1, kGroup | kNoPar)
    SQE_PHASE_ARM(kArmPhaseNp + 2, kHalf + 0 + kNoPar)
    SQE_PHASE_ARM(kArmPhaseNp + 3, kHalf + 1 + kNoPar)
    SQE_PHASE_ARM(kArmPhaseNp + 4, kHalf + 2 + kNoPar)
    SQE_PHASE_ARM(kArmPhaseNp + 5, kHalf + 3 + kNoPar)
    SQE_PHASE_ARM(kArmPhaseNp + 6, kHalf + 4 + kNoPar)
    SQE_PHASE_ARM(kArmPhaseNp + 7, kHalf + 5 + kNoPar)
    SQE_NP_ARMS(kPauliR, 0)
    SQE_NP_ARMS(kPauliI, 1)
    default: return x;   // never set by the host
    }
}
#undef SQE_EACH
#undef SQE_PAIR_ARMS
#undef SQE_MULTI_ARMS
#undef SQE_SWAP_HALF
#undef SQE_PHASE_ARM
#undef SQE_GROUP_ARMS
#undef SQE_NP_ARMS

// The register-step part of a gate descriptor, loaded whole (FP32 passes skip
// the FP64 coefficients).
typedef uint32_t u32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
template <typename T>
__device__ __forceinline__ GateR load_gate(const TileGate& tg) {
    GateR r;
    const u32x16 h = *reinterpret_cast<const u32x16*>(&tg);
    __builtin_memcpy(&r, &h, 64);
    if constexpr (std::is_same<T, double>::value) {
        const u32x4 h2 = *reinterpret_cast<const u32x4*>(reinterpret_cast<const char*>(&tg) + 64);
        __builtin_memcpy(reinterpret_cast<char*>(&r) + 64, &h2, 16);
    } else {
        r.c = r.s = 0.0;
    }
    return r;
}

// deposit(x, mask) for a mask given as runs of consecutive bits (host-built):
// a few instructions per run instead of per mask bit.
__device__ __forceinline__ uint64_t deposit_runs(uint64_t x, const uint32_t* runs, uint32_t n) {
    uint64_t r = 0;
#pragma unroll
    for (uint32_t i = 0; i < kMaxRuns; ++i) {
        if (i < n) {
            const uint32_t w = runs[i];
            const uint32_t lo = w & 63u, len = (w >> 8) & 63u, src = w >> 16;
            r |= ((x >> src) & ((1ull << len) - 1ull)) << lo;
        }
    }
    return r;
}

template <typename T, int G>
__global__ __launch_bounds__(kBlock) void k_tile(T* __restrict__ re, T* __restrict__ im, const TileArgs a) {
    typedef typename AmpT<T>::chunk CH;
    constexpr uint32_t V = AmpT<T>::kVec;
    constexpr uint32_t kMaxChunks = (1u << AmpT<T>::kMaxBits) / V / kBlock;   // chunks per array per thread
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const uint32_t tsz = 1u << a.t;
    Cplx<T>* lc = reinterpret_cast<Cplx<T>*>(lds_raw);
    const uint64_t NB = gridDim.x, b = blockIdx.x;
    const uint64_t tile = (b & 7ull) * (NB >> 3) + (b >> 3);   // XCD-contiguous (grid % 8 == 0)
    if (tile >= a.ntiles) return;                             // whole workgroup leaves together
    const uint64_t base = deposit_runs(tile, a.free_runs, a.nfree_runs) | a.fixval;

    // 1. tile -> LDS: chunk k of a thread is tile elements e = V tid + k V kBlock
    //    onwards (tile bits 0..log2 V - 1 are state qubits 0.. of one line); every
    //    load of the thread is issued before the LDS writes
    const uint32_t e0 = V * threadIdx.x;
    const uint32_t nch = tsz / (V * kBlock);   // 0 when the tile is smaller than V kBlock
    uint64_t toff = 0;
    if (e0 < tsz) {
        if (a.skew) {   // base is zero at the vectors' pivots: XOR, not OR
            constexpr uint32_t lv = V == 4 ? 2u : 1u;   // tile bit of thread index bit 0
            uint64_t l = 0;
#pragma unroll
            for (uint32_t m = 0; m < 8; ++m) l ^= ((threadIdx.x >> m) & 1u) ? a.lvec[m + lv] : 0ull;
            toff = base ^ l;
        } else {
            toff = base | deposit_runs(e0, a.tile_runs, a.ntile_runs);
        }
    }
    // state offsets of the tile bits that chunk index bits 0, 1, 2 stand for
    const uint64_t* kbit = a.kbit;
    auto chunk_off = [&](uint32_t k) {
        return ((k & 1u) ? kbit[0] : 0ull) ^ ((k & 2u) ? kbit[1] : 0ull) ^ ((k & 4u) ? kbit[2] : 0ull);
    };
    if (nch == 0) {
        if (e0 < tsz) {
            const CH vr = __builtin_nontemporal_load(reinterpret_cast<const CH*>(re + toff));
            const CH vi = __builtin_nontemporal_load(reinterpret_cast<const CH*>(im + toff));
#pragma unroll
            for (uint32_t i = 0; i < V; ++i) lc[e0 + i] = Cplx<T>{vr[i], vi[i]};
        }
    } else {
        CH vr[kMaxChunks], vi[kMaxChunks];
#pragma unroll
        for (uint32_t k = 0; k < kMaxChunks; ++k) {
            if (k < nch) {
                const uint64_t g = toff ^ chunk_off(k);
                vr[k] = __builtin_nontemporal_load(reinterpret_cast<const CH*>(re + g));
                vi[k] = __builtin_nontemporal_load(reinterpret_cast<const CH*>(im + g));
            }
        }
#pragma unroll
        for (uint32_t k = 0; k < kMaxChunks; ++k) {
            if (k < nch) {
#pragma unroll
                for (uint32_t i = 0; i < V; ++i) lc[e0 + k * V * kBlock + i] = Cplx<T>{vr[k][i], vi[k][i]};
            }
        }
    }
    __syncthreads();

    // 2. the steps, in order
    for (uint32_t st = 0; st < a.nsteps; ++st) {
        const auto& S = a.step[st];
        if (S.f[0] == kLdsStep) {   // one gate flipping more than kSub bits: pairs in LDS
            const TileGate& g = a.g[S.g0];
            if ((base & g.mo) == g.vo) {
                const uint32_t opar = (uint32_t)__popcll(base & g.yzo) & 1u;
                switch (g.kind) {
                case OP_SWAP: tile_pairs<OP_SWAP>(lc, tsz / 2, g, opar); break;
                case OP_Y: tile_pairs<OP_Y>(lc, tsz / 2, g, opar); break;
                case OP_H: tile_pairs<OP_H>(lc, tsz / 2, g, opar); break;
                default: tile_pairs<OP_PAULI>(lc, tsz / 2, g, opar); break;
                }
            }
            __syncthreads();
            continue;
        }
        uint32_t off[kLoc];   // uniform: element j of a group sits at gb + off[j]
#pragma unroll
        for (uint32_t j = 0; j < kLoc; ++j)
            off[j] = ((j & 1u) << S.f[0]) | (((j >> 1) & 1u) << S.f[1]) | (((j >> 2) & 1u) << S.f[2]);
        // G groups per thread at a time (the host picks G with tsz / kLoc a multiple of G kBlock)
        for (uint32_t gi0 = threadIdx.x; gi0 < tsz / kLoc; gi0 += G * kBlock) {
            uint32_t gb[G];
            RMG<T, G> x;
#pragma unroll
            for (int q = 0; q < G; ++q) {
                uint32_t b2 = gi0 + (uint32_t)q * kBlock;   // insert zeros at the step's bits (ascending)
#pragma unroll
                for (uint32_t z = 0; z < kSub; ++z) {
                    const uint32_t lo = (1u << S.f[z]) - 1u;
                    b2 = (b2 & lo) | ((b2 & ~lo) << 1);
                }
                gb[q] = b2;
                const Cplx<T>* pg = lc + b2;   // gb has zeros at the step's bits: gb | off = gb + off
#pragma unroll
                for (uint32_t j = 0; j < kLoc; ++j) x.q[q].e[j] = pg[off[j]];
            }
            for (uint32_t k = S.g0; k < S.g1; ++k) {
                const GateR g = load_gate<T>(a.g[k]);
                if ((base & g.mo) != g.vo) continue;   // uniform: tile outside the gate's selection
                const uint32_t opar = (uint32_t)__popcll(base & g.yzo);
                GroupSel<G> p;
#pragma unroll
                for (int q = 0; q < G; ++q) {
                    p.pass[q] = (gb[q] & g.mcg) == g.vcg;
                    p.par[q] = (opar + (uint32_t)__popc(gb[q] & g.yzg)) & 1u;
                }
                x = reg_gate<T, G>(x, g, p);
            }
#pragma unroll
            for (int q = 0; q < G; ++q) {
                Cplx<T>* pg = lc + gb[q];
#pragma unroll
                for (uint32_t j = 0; j < kLoc; ++j) pg[off[j]] = x.q[q].e[j];
            }
        }
        __syncthreads();
    }

    // 3. LDS -> tile
    if (nch == 0) {
        if (e0 < tsz) {
            CH vr, vi;
#pragma unroll
            for (uint32_t i = 0; i < V; ++i) {
                const Cplx<T> v = lc[e0 + i];
                vr[i] = v.x;
                vi[i] = v.y;
            }
            __builtin_nontemporal_store(vr, reinterpret_cast<CH*>(re + toff));
            __builtin_nontemporal_store(vi, reinterpret_cast<CH*>(im + toff));
        }
    } else {
#pragma unroll
        for (uint32_t k = 0; k < kMaxChunks; ++k) {
            if (k < nch) {
                const uint64_t g = toff ^ chunk_off(k);
                CH vr, vi;
#pragma unroll
                for (uint32_t i = 0; i < V; ++i) {
                    const Cplx<T> v = lc[e0 + k * V * kBlock + i];
                    vr[i] = v.x;
                    vi[i] = v.y;
                }
                __builtin_nontemporal_store(vr, reinterpret_cast<CH*>(re + g));
                __builtin_nontemporal_store(vi, reinterpret_cast<CH*>(im + g));
            }
        }
    }
}

// A pass over a state that is one tile of at most 2^kTinyBits amplitudes: a
// workgroup of 2^t threads, one amplitude per thread.  A pair gate exchanges
// each amplitude with its partner lane (a shuffle inside the wavefront, LDS
// across wavefronts) and every thread computes its own outputs with the
// arithmetic and roles of tile_pairs; a phase gate is element-wise as in
// tile_phase.  Latency-bound states take one exchange per gate instead of
// register steps on tsz / 8 threads.  The pass's gate descriptors are copied
// from the kernel arguments into LDS by all threads at once first: read gate by
// gate from the argument segment, each gate waited for its own scalar loads
// (about half a microsecond per gate).
__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
    return ((uint64_t)uni((uint32_t)(v >> 32)) << 32) | uni((uint32_t)v);
}
__device__ __forceinline__ double unid(double v) {
    return __longlong_as_double((long long)uni64((uint64_t)__double_as_longlong(v)));
}
__device__ __forceinline__ float unif(float v) { return __uint_as_float(uni(__float_as_uint(v))); }

template <typename T>
__global__ __launch_bounds__(1024) void k_tile_tiny(T* __restrict__ re, T* __restrict__ im, const TileArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    T* lds = reinterpret_cast<T*>(lds_raw);
    const uint32_t tsz = 1u << a.t;
    const uint32_t y = threadIdx.x;
    const uint64_t base = a.fixval;   // one tile: every other state bit is fixed by the pass
    const uint64_t off = base | deposit(y, a.tmask);
    T vr = re[off], vi = im[off];
    TileGate* sg = reinterpret_cast<TileGate*>(lds + 4 * (size_t)tsz);   // after the exchange buffers
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(&a.g[0]);
        uint32_t* dst = reinterpret_cast<uint32_t*>(sg);
        const uint32_t nw = a.ngates * (uint32_t)(sizeof(TileGate) / 4);
        for (uint32_t i = y; i < nw; i += tsz) dst[i] = src[i];
    }
    __syncthreads();
    uint32_t xb = 0;   // alternating exchange buffers: one barrier per exchange
    for (uint32_t k = 0; k < a.ngates; ++k) {
        TileGate g;   // the fields used here, made uniform (every lane read the same LDS words)
        g.kind = uni(sg[k].kind);
        g.pivot = uni(sg[k].pivot);
        g.xc = uni(sg[k].xc);
        g.mc = uni(sg[k].mc);
        g.vc = uni(sg[k].vc);
        g.yzc = uni(sg[k].yzc);
        g.base_imag = uni(sg[k].base_imag);
        g.base_neg = uni(sg[k].base_neg);
        g.mo = uni64(sg[k].mo);
        g.vo = uni64(sg[k].vo);
        g.yzo = uni64(sg[k].yzo);
        if constexpr (std::is_same<T, double>::value) {
            g.c = unid(sg[k].c);
            g.s = unid(sg[k].s);
        } else {
            g.cf = unif(sg[k].cf);
            g.sf = unif(sg[k].sf);
        }
        if ((base & g.mo) != g.vo) continue;   // uniform: the tile is outside the gate's selection
        const uint32_t opar = (uint32_t)__popcll(base & g.yzo) & 1u;
        if (g.kind == (uint32_t)kPhaseOp) {
            if ((y & g.mc) == g.vc) {
                const uint32_t par = ((uint32_t)__popc(y & g.yzc) & 1u) ^ opar;
                const T s = gate_s<T>(g);
                const T sn = par ? -s : s;
                const T rr = vr, mm = vi;
                vr = phase_out(false, rr, mm, gate_c<T>(g), sn);
                vi = phase_out(true, mm, rr, gate_c<T>(g), sn);
            }
            continue;
        }
        const uint32_t w = y ^ g.xc;   // every thread takes part in the exchange
        T pr, pi;
        if (g.xc < 64u) {
            pr = __shfl_xor(vr, (int)g.xc, 64);
            pi = __shfl_xor(vi, (int)g.xc, 64);
        } else {   // the previous use of this buffer ended before the last exchange's barrier
            T* bx = lds + 2 * tsz * xb;
            bx[y] = vr;
            bx[tsz + y] = vi;
            __syncthreads();
            pr = bx[w];
            pi = bx[tsz + w];
            xb ^= 1u;
        }
        // roles of tile_pairs: u has the pivot bit 0; the selected one of (u, u ^ X) is the a side
        const uint32_t u = ((y >> g.pivot) & 1u) ? w : y;
        const uint32_t uw = u ^ g.xc;
        const bool su = (u & g.mc) == g.vc;
        if (!su && (uw & g.mc) != g.vc) continue;
        const bool is_a = y == (su ? u : uw);
        const uint32_t par = ((uint32_t)__popc(w & g.yzc) & 1u) ^ opar;   // parity of the partner
        const Coef<T> pc = coef<T>(g);
        T nr, ni;
        switch (g.kind) {
        case OP_SWAP:
            nr = pair_out_par<OP_SWAP>(is_a, false, vr, pr, pi, par, pc);
            ni = pair_out_par<OP_SWAP>(is_a, true, vi, pi, pr, par, pc);
            break;
        case OP_Y:
            nr = pair_out_par<OP_Y>(is_a, false, vr, pr, pi, par, pc);
            ni = pair_out_par<OP_Y>(is_a, true, vi, pi, pr, par, pc);
            break;
        case OP_H:
            nr = pair_out_par<OP_H>(is_a, false, vr, pr, pi, par, pc);
            ni = pair_out_par<OP_H>(is_a, true, vi, pi, pr, par, pc);
            break;
        default:
            nr = pair_out_par<OP_PAULI>(is_a, false, vr, pr, pi, par, pc);
            ni = pair_out_par<OP_PAULI>(is_a, true, vi, pi, pr, par, pc);
            break;
        }
        vr = nr;
        vi = ni;
    }
    re[off] = vr;
    im[off] = vi;
}

// ------------------------------------------------------------ tiny programs
// A queued operation of a state of at most 2^kTinyProgBits amplitudes, as
// k_tiny_prog reads it (40 bytes).  The state is one tile: every mask is over
// the state's own bits.
struct ProgOp {
    uint32_t kf;    // bits 0-3 kind; 4 base_imag; 5 base_neg; 6 reset; 7 the refs are op indices of
                    // this program (else outcome slots); 8 conditional gate; 9-13 pivot; 16-23 E_old bits
    uint32_t x;     // pair: partner mask; measurement: the qubit's bit; W: E_new mask
    uint32_t m, v;  // selection (gates); measurement: gc bits in m; W: E_old mask in m, b0 | ob0 << 8 in v
    uint32_t yz;    // parity mask (gates); measurement: its slot; W: L
    uint32_t ref;   // conditional gate: the outcome it reads; W: E_old's first outcome, or E_old itself
    double c, s;    // gate coefficients; measurement: the draw in c; W: the table's address in c
};
static_assert(sizeof(ProgOp) == 40, "ProgOp layout");
constexpr uint32_t kPfImag = 1u << 4, kPfNeg = 1u << 5, kPfReset = 1u << 6, kPfLocal = 1u << 7, kPfCond = 1u << 8;
struct ProgArgs {
    double* re;
    double* im;
    uint32_t* slots;   // the outcome ring
    uint32_t nops, nq;
    ProgOp op[kProgMax];
};
static_assert(sizeof(ProgArgs) <= 4096, "kernel arguments are limited to 4 KiB");

// One workgroup of max(N, kBlock) threads, amplitude y in thread y < N, the
// program's operations in order: gates as k_tile_tiny (bit-identical to one
// sweep per gate), measurements as k_measure_small (the same partial sums in
// the same order: thread t < kBlock sums amplitudes t, t + kBlock, ...; the
// wave butterfly; wave 0..3 in order), apply_W as k_qms_W.
__global__ __launch_bounds__(1024) void k_tiny_prog(const ProgArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ __align__(16) ProgOp sop[kProgMax];
    __shared__ uint32_t lout[kProgMax];          // outcome of each measurement op of this program
    __shared__ double red[2][2][kBlock / 64];    // alternating wave partials: one barrier per measurement
    const uint32_t N = 1u << a.nq;
    const uint32_t y = threadIdx.x;
    const bool act = y < N;
    double* xbuf = reinterpret_cast<double*>(lds_raw);   // 2 exchange buffers of (re[N], im[N])
    double* pbuf = xbuf + 4 * (size_t)N;                 // |psi|^2 for the measurements (N > kBlock)
    double vr = act ? a.re[y] : 0.0, vi = act ? a.im[y] : 0.0;
    {   // the ops into LDS, 16 bytes per thread (kProgMax * 40 / 16 <= kBlock <= blockDim.x)
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
        const uint32_t n16 = (a.nops * (uint32_t)sizeof(ProgOp) + 15u) / 16u;
        if (y < n16) reinterpret_cast<u4*>(sop)[y] = reinterpret_cast<const u4*>(&a.op[0])[y];
    }
    __syncthreads();
    uint32_t xb = 0;   // alternating exchange buffers: one barrier per cross-wave exchange
    uint32_t rb = 0;   // alternating reduction partials
    // partner amplitude y ^ x (every thread calls it: barriers inside)
    auto exchange = [&](uint32_t x, double& pr, double& pi) {
        if (x < 64u) {
            pr = __shfl_xor(vr, (int)x, 64);
            pi = __shfl_xor(vi, (int)x, 64);
        } else {   // the previous use of this buffer ended before the last exchange's barrier
            double* bx = xbuf + 2 * (size_t)N * xb;
            if (act) {
                bx[y] = vr;
                bx[N + y] = vi;
            }
            __syncthreads();
            if (act) {
                pr = bx[y ^ x];
                pi = bx[N + (y ^ x)];
            } else {
                pr = pi = 0.0;
            }
            xb ^= 1u;
        }
    };
    // outcome k of an op's references (uniform)
    auto outcome = [&](uint32_t kf, uint32_t ref, uint32_t k) -> uint32_t {
        if (kf & kPfLocal) return uni(lout[ref + k]);
        return uni(__hip_atomic_load(a.slots + ((ref + k) % kSlots), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    };
    // the next op's words are read from LDS while the current one runs
    ProgOp nx = sop[0];
    for (uint32_t k = 0; k < a.nops; ++k) {
        const uint32_t kf = uni(nx.kf);
        const uint32_t kind = kf & 15u;
        const uint32_t x = uni(nx.x), m = uni(nx.m), v = uni(nx.v), yz = uni(nx.yz);
        const uint32_t ref = uni(nx.ref);
        const double c = unid(nx.c), s = unid(nx.s);
        if (k + 1 < a.nops) nx = sop[k + 1];
        if (kind == (uint32_t)kMeasOp) {   // k_measure_small's sums for one qubit
            double s0 = 0.0, s1 = 0.0;
            if (N <= (uint32_t)kBlock) {   // one term per thread
                const double p = vr * vr + vi * vi;
                if (act) {
                    if (y & x) s1 += p;
                    else s0 += p;
                }
            } else {   // thread t < kBlock sums amplitudes t, t + kBlock, ...
                if (act) pbuf[y] = vr * vr + vi * vi;
                __syncthreads();
                if (y < (uint32_t)kBlock)
                    for (uint32_t i = y; i < N; i += kBlock) {
                        const double p = pbuf[i];
                        if (i & x) s1 += p;
                        else s0 += p;
                    }
            }
            if (y < (uint32_t)kBlock) {
                s0 = wave_sum64(s0);
                s1 = wave_sum64(s1);
                if ((y & 63u) == 0u) {
                    red[rb][0][y >> 6] = s0;
                    red[rb][1][y >> 6] = s1;
                }
            }
            __syncthreads();
            double p0 = 0.0, p1 = 0.0;
            for (int w = 0; w < kBlock / 64; ++w) {
                p0 += red[rb][0][w];
                p1 += red[rb][1][w];
            }
            rb ^= 1u;
            const uint32_t cbit = uni(c < p1 ? 1u : 0u);        // suqa.cu:690
            const double scale = 1. / sqrt(cbit ? p1 : p0);    // set_ampl_to_zero + vnormalize (suqa.cu:694-695)
            const uint32_t keep = cbit ? x : 0u;
            if ((y & 63u) == 0u) lout[k] = cbit;   // each wave's own write precedes its reads
            if (y == 0) a.slots[yz] = cbit;
            if (act) {
                const bool kp = (y & x) == keep;
                vr = kp ? vr * scale : 0.0;
                vi = kp ? vi * scale : 0.0;
            }
            if ((kf & kPfReset) && cbit) {   // apply_reset: X on q over the gc bits (suqa.cu:735-737)
                double pr, pi;
                exchange(x, pr, pi);
                const uint32_t g = m & ~x;
                if (act && (y & g) == g) {
                    vr = pr;
                    vi = pi;
                }
            }
            continue;
        }
        if (kind == (uint32_t)kWOp) {   // k_qms_W: pairs (j, i = j + N/2)
            const uint32_t half = N >> 1;
            double pr, pi;
            exchange(half, pr, pi);
            if (act) {
                const uint32_t i = y | half;
                const uint32_t En = (i & x) >> (v & 255u);
                uint32_t Eo;
                const uint32_t eb = (kf >> 16) & 255u;
                if (m) {
                    Eo = (i & m) >> ((v >> 8) & 255u);
                } else if (eb) {
                    Eo = 0;
                    for (uint32_t b = 0; b < eb; ++b) Eo |= outcome(kf, ref, b) << b;
                } else {
                    Eo = ref;
                }
                const double* row = reinterpret_cast<const double*>(__double_as_longlong(c)) + 2 * (uint64_t)Eo * yz;
                const double fs1 = row[2 * En], fs2 = row[2 * En + 1];
                if (y < half) {   // j: t = own, ri = the partner
                    vr = fs2 * vr + fs1 * pr;
                    vi = fs2 * vi + fs1 * pi;
                } else {          // i: t = the partner, ri = own
                    vr = fs1 * pr - fs2 * vr;
                    vi = fs1 * pi - fs2 * vi;
                }
            }
            continue;
        }
        if ((kf & kPfCond) && !outcome(kf, ref, 0)) continue;   // uniform: a conditional gate not taken
        if (kind == (uint32_t)kPhaseOp) {
            if (act && (y & m) == v) {
                const uint32_t par = (uint32_t)__popc(y & yz) & 1u;
                const double sn = par ? -s : s;
                const double rr = vr, mm = vi;
                vr = phase_out(false, rr, mm, c, sn);
                vi = phase_out(true, mm, rr, c, sn);
            }
            continue;
        }
        double pr, pi;
        exchange(x, pr, pi);
        if (!act) continue;
        // roles of tile_pairs: u has the pivot bit 0; the selected one of (u, u ^ X) is the a side
        const uint32_t w = y ^ x;
        const uint32_t pivot = (kf >> 9) & 31u;
        const uint32_t u = ((y >> pivot) & 1u) ? w : y;
        const uint32_t uw = u ^ x;
        const bool su = (u & m) == v;
        if (!su && (uw & m) != v) continue;
        const bool is_a = y == (su ? u : uw);
        const uint32_t par = (uint32_t)__popc(w & yz) & 1u;   // parity of the partner
        const Coef<double> pc{c, s, (kf & kPfImag) ? 1u : 0u, (kf & kPfNeg) ? 1u : 0u};
        double nr, ni;
        switch (kind) {
        case OP_SWAP:
            nr = pair_out_par<OP_SWAP>(is_a, false, vr, pr, pi, par, pc);
            ni = pair_out_par<OP_SWAP>(is_a, true, vi, pi, pr, par, pc);
            break;
        case OP_Y:
            nr = pair_out_par<OP_Y>(is_a, false, vr, pr, pi, par, pc);
            ni = pair_out_par<OP_Y>(is_a, true, vi, pi, pr, par, pc);
            break;
        case OP_H:
            nr = pair_out_par<OP_H>(is_a, false, vr, pr, pi, par, pc);
            ni = pair_out_par<OP_H>(is_a, true, vi, pi, pr, par, pc);
            break;
        default:
            nr = pair_out_par<OP_PAULI>(is_a, false, vr, pr, pi, par, pc);
            ni = pair_out_par<OP_PAULI>(is_a, true, vi, pi, pr, par, pc);
            break;
        }
        vr = nr;
        vi = ni;
    }
    if (act) {
        a.re[y] = vr;
        a.im[y] = vi;
    }
}

// Runs of consecutive set bits of m, lowest first, as lo | len << 8 | src << 16
// (src: the index of the run's first bit among m's set bits); only the runs
// whose src is below src_limit.  At most kMaxRuns (a 64-bit mask of t tile
// bits has at most t + 1 runs of clear bits and t of set bits, t <= 13).
uint32_t mask_runs(uint64_t m, uint32_t* runs, uint32_t src_limit) {
    uint32_t n = 0, src = 0;
    while (m && src < src_limit && n < kMaxRuns) {
        const uint32_t lo = (uint32_t)__builtin_ctzll(m);
        const uint64_t shifted = m >> lo;
        const uint32_t len = shifted == ~0ull ? 64u - lo : (uint32_t)__builtin_ctzll(~shifted);
        runs[n++] = lo | (len << 8) | (src << 16);
        src += len;
        m &= len >= 64 ? 0ull : ~(((1ull << len) - 1ull) << lo);
    }
    return n;
}

hipError_t run_single(suqa_hip_ctx* c, const GateDesc& d) {
    Sel sel;
    if (!make_sel(c->nq, d.M, d.V, &sel)) return hipSuccess;
    ++c->single_sweeps;
    if (d.kind == kPhaseOp) return launch_phase_now(c, sel, d.c, d.s, d.zmask, d.tag, d.alg_bytes);
    return launch_pair_now(c, d.kind, sel, d.X, d.pc, d.tag, d.alg_bytes);
}

// Local tables of gate g for a register step on tile positions f[0..2].
void local_tables(TileGate& g, const uint32_t* f) {
    uint32_t F = 0, mcl = 0, vcl = 0, yzl = 0, xl = 0;
    for (uint32_t m = 0; m < kSub; ++m) {
        F |= 1u << f[m];
        if ((g.mc >> f[m]) & 1u) mcl |= 1u << m;
        if ((g.vc >> f[m]) & 1u) vcl |= 1u << m;
        if ((g.yzc >> f[m]) & 1u) yzl |= 1u << m;
        if ((g.xc >> f[m]) & 1u) xl |= 1u << m;
    }
    g.xl = xl;
    g.lsel = 0;
    g.lpar = 0;
    for (uint32_t j = 0; j < kLoc; ++j) {
        if ((j & mcl) == vcl) g.lsel |= 1u << j;
        if (__builtin_popcount(j & yzl) & 1) g.lpar |= 1u << j;
    }
    g.mcg = g.mc & ~F;
    g.vcg = g.vc & ~F;
    g.yzg = g.yzc & ~F;
    g.gtriv = g.mcg == 0;
    // dense: no selection left to test.  Phase: all 8 elements selected; one-bit
    // pair: all 4 lower elements (pivot bit 0) selected
    uint32_t need = 0xffu;
    if (g.kind != (uint32_t)kPhaseOp) {
        need = 0;
        if (__builtin_popcount(xl) == 1)
            for (uint32_t j = 0; j < kLoc; ++j)
                if ((j & xl) == 0) need |= 1u << j;
    }
    g.dense = g.gtriv && need != 0 && (g.lsel & need) == need;
    // the code path of reg_gate
    const bool one_local_sel = g.gtriv && __builtin_popcount(mcl) == 1;   // selection: one F bit, nothing else
    const bool nopar = g.yzc == 0u && g.yzo == 0ull;   // no parity qubit anywhere: a uniform sign
    if (g.kind == (uint32_t)kPhaseOp) {
        if (g.dense) g.arm = nopar ? kArmPhaseNp : kArmPhaseDense;
        else if (!g.gtriv && mcl == 0u) g.arm = nopar ? kArmPhaseNp + 1u : kArmPhaseGroup;
        else if (one_local_sel) {
            const uint32_t b = (uint32_t)__builtin_ctz(mcl);
            g.arm = (nopar ? kArmPhaseNp + 2u : kArmPhaseHalf) + 2u * b + ((vcl >> b) & 1u);
        } else g.arm = kArmPhaseGeneral;
        return;
    }
    const uint32_t opv = g.kind == (uint32_t)OP_PAULI ? (g.base_imag ? (uint32_t)kPauliI : (uint32_t)kPauliR) : g.kind;
    const uint32_t base = kArmPair + 16u * opv;
    if (__builtin_popcount(xl) == 1) {
        const uint32_t lx = (uint32_t)__builtin_ctz(xl);
        const uint32_t ctl = mcl & ~xl;   // selection bits on F besides the pivot (selected at 0)
        const bool pauli = g.kind == (uint32_t)OP_PAULI;
        if (g.dense) g.arm = pauli && nopar ? kArmPauliNpDense + 3u * (opv - kPauliR) + lx : base + kArmPairDense + lx;
        else if (!g.gtriv && ctl == 0u)
            g.arm = pauli && nopar ? kArmPauliNpGroup + 3u * (opv - kPauliR) + lx : kArmPairGroup + 3u * opv + lx;
        else if (g.kind == (uint32_t)OP_SWAP && g.gtriv && __builtin_popcount(ctl) == 1) {
            const uint32_t cb = (uint32_t)__builtin_ctz(ctl);
            const uint32_t k = cb > lx ? cb - 1u : cb;   // index of cb among the two other bits
            g.arm = kArmSwapHalf + 4u * lx + 2u * (k == 0u ? 0u : 1u) + ((vcl >> cb) & 1u);
        } else g.arm = base + kArmPairGeneral1 + lx;
    } else {   // SWAP and Pauli only (run_pass sends other multi-bit gates to LDS steps)
        g.arm = base + kArmPairMulti + (xl == 3u ? 0u : xl == 5u ? 1u : xl == 6u ? 2u : 3u);
    }
}

uint64_t gate_flips(const GateDesc& d) { return d.kind == kPhaseOp ? 0ull : d.X; }

// The tile of a pass.  Tile coordinate m stands for the state vector vec[m]:
// element e of the tile with base b is b ^ (XOR of vec[m] over e's set bits m).
// Coordinates 0..f-1 are the low qubits; the others are, ordered by pivot, a
// unit vector for every selection qubit inside the flip masks' support, the
// linearly independent flip masks (above the low qubits, those selection
// qubits cleared; span tiles only), and unit vectors that pad to t
// coordinates.  With span tiles off every flipped qubit is a unit vector.
// Gaussian elimination in coordinate order gives every vector a pivot bit
// (its lowest bit once reduced by the earlier rows at their pivots); tile
// bases are zero at the pivots, and the vectors restricted to the pivots form
// an invertible matrix, so the tiles partition the state.  A selection qubit
// in some vector's support is a coordinate only when its unit vector is the one
// vector holding it, which the construction guarantees.
struct TileBasis {
    uint32_t t = 0;
    uint64_t vec[kMaxTileBits32] = {};
    uint64_t P = 0;        // pivot bits
    uint64_t TB = 0;       // union of the vectors' supports
    bool skew = false;     // some vector holds more than one qubit
    uint64_t row[kMaxTileBits32] = {};   // elimination rows: row[i] = XOR of vec[m] over comb[i]'s bits
    uint32_t comb[kMaxTileBits32] = {};
    uint32_t rp[kMaxTileBits32] = {};    // row pivots
    // coordinates of a state vector in the tile, false when it is outside the span
    bool coords(uint64_t v, uint32_t* out) const {
        uint32_t cm = 0;
        for (uint32_t i = 0; i < t; ++i)
            if ((v >> rp[i]) & 1ull) {
                v ^= row[i];
                cm ^= comb[i];
            }
        *out = cm;
        return v == 0ull;
    }
    // the coordinate that reads qubit b alone, -1 when none does
    int unit_coord(uint32_t b) const {
        int m0 = -1;
        for (uint32_t m = 0; m < t; ++m)
            if ((vec[m] >> b) & 1ull) {
                if (m0 >= 0 || vec[m] != (1ull << b)) return -1;
                m0 = (int)m;
            }
        return m0;
    }
};

bool make_basis(const GateDesc* q, uint32_t n, uint32_t t, uint32_t f, uint32_t nq, bool span, TileBasis& B) {
    const uint64_t low = (1ull << f) - 1ull;
    uint64_t S = 0, supp = 0;
    for (uint32_t i = 0; i < n; ++i) {
        S |= pass_sel(q[i]) & ~low;
        supp |= gate_flips(q[i]) & ~low;
    }
    const uint64_t Q = span ? (S & supp) : supp;
    struct HV {
        uint64_t v;
        uint32_t pivot;
    };
    std::vector<HV> hv;
    for (uint64_t m = Q; m; m &= m - 1) hv.push_back(HV{m & (0ull - m), (uint32_t)__builtin_ctzll(m)});
    uint64_t hpiv = 0;
    if (span) {   // independent flip masks, kept whole; pivots by elimination
        std::vector<uint64_t> red;
        std::vector<uint32_t> rpv;
        for (uint32_t i = 0; i < n; ++i) {
            const uint64_t h = gate_flips(q[i]) & ~low & ~Q;
            if (!h) continue;
            uint64_t x = h;
            for (size_t j = 0; j < red.size(); ++j)
                if ((x >> rpv[j]) & 1ull) x ^= red[j];
            if (!x) continue;   // in the span already
            const uint32_t p = (uint32_t)__builtin_ctzll(x);
            red.push_back(x);
            rpv.push_back(p);
            hpiv |= 1ull << p;
            hv.push_back(HV{h, p});
        }
    }
    if (hv.size() + f > t) return false;
    uint32_t need = t - f - (uint32_t)hv.size();
    for (uint32_t b = f; b < nq && need; ++b)   // padding: qubits outside the support first
        if (!((supp >> b) & 1ull)) {
            hv.push_back(HV{1ull << b, b});
            --need;
        }
    for (uint32_t b = f; b < nq && need; ++b)   // then support qubits that are neither pivots nor selected
        if (((supp >> b) & 1ull) && !(((Q | hpiv) >> b) & 1ull)) {
            hv.push_back(HV{1ull << b, b});
            --need;
        }
    if (need) return false;
    std::stable_sort(hv.begin(), hv.end(), [](const HV& x, const HV& y) { return x.pivot < y.pivot; });
    B.t = t;
    for (uint32_t m = 0; m < f; ++m) B.vec[m] = 1ull << m;
    for (size_t k = 0; k < hv.size(); ++k) B.vec[f + k] = hv[k].v;
    B.P = B.TB = 0;
    B.skew = false;
    for (uint32_t i = 0; i < t; ++i) {
        uint64_t x = B.vec[i];
        uint32_t cm = 1u << i;
        for (uint32_t j = 0; j < i; ++j)
            if ((x >> B.rp[j]) & 1ull) {
                x ^= B.row[j];
                cm ^= B.comb[j];
            }
        if (!x) return false;   // dependent vectors: a planning error
        B.row[i] = x;
        B.comb[i] = cm;
        B.rp[i] = (uint32_t)__builtin_ctzll(x);
        B.P |= 1ull << B.rp[i];
        B.TB |= B.vec[i];
        if (B.vec[i] & (B.vec[i] - 1ull)) B.skew = true;
    }
    return true;
}

hipError_t run_pass(suqa_hip_ctx* c, const GateDesc* q, uint32_t n, uint32_t t, uint32_t f, bool span) {
    const uint32_t nq = c->nq;
    TileBasis B;
    if (!make_basis(q, n, t, f, nq, span, B)) {
        set_last_error("pass tile: the gates' flip masks do not fit the tile");
        return hipErrorInvalidValue;
    }
    TileArgs a;
    std::memset(&a, 0, sizeof(a));
    a.t = t;
    a.tmask = B.P;
    a.skew = B.skew ? 1u : 0u;
    for (uint32_t m = 0; m < t; ++m) a.lvec[m] = B.vec[m];
    a.ngates = n;
    uint64_t mcom = ~0ull, vcom = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const GateDesc& d = q[i];
        TileGate& g = a.g[i];
        g.kind = (uint32_t)d.kind;
        uint32_t xcoord = 0;
        if (d.kind != kPhaseOp && (!B.coords(d.X, &xcoord) || xcoord == 0u)) {
            set_last_error("pass tile: a flip mask outside the tile");
            return hipErrorInvalidValue;
        }
        g.xc = (uint16_t)xcoord;
        // selection: coordinates for qubits the tile reads, tile-level bits for
        // the rest; a Pauli string's pair enumeration qubit (pass_sel) becomes
        // its lowest flipped coordinate at 0
        const uint64_t msel = pass_sel(d);
        uint32_t mc = msel != d.M ? (g.xc & (0u - g.xc)) : 0u, vc = 0;
        uint64_t mo = 0, vo = 0;
        for (uint64_t mm = msel; mm; mm &= mm - 1) {
            const uint32_t b = (uint32_t)__builtin_ctzll(mm);
            const uint32_t vb = (uint32_t)((d.V >> b) & 1ull);
            if ((B.TB >> b) & 1ull) {
                const int m = B.unit_coord(b);
                if (m < 0) {
                    set_last_error("pass tile: a selection qubit is not a tile coordinate");
                    return hipErrorInvalidValue;
                }
                mc |= 1u << m;
                vc |= vb << m;
            } else {
                mo |= 1ull << b;
                vo |= (uint64_t)vb << b;
            }
        }
        g.mc = mc;
        g.vc = vc;
        g.mo = mo;
        g.vo = vo;
        const uint64_t par = d.kind == kPhaseOp ? d.zmask : d.pc.yz;
        uint32_t yzc = 0;
        for (uint32_t m = 0; m < t; ++m) yzc |= (uint32_t)(__builtin_popcountll(B.vec[m] & par) & 1) << m;
        g.yzc = yzc;
        g.yzo = par & ~B.P;   // bases are zero at the pivots
        if (d.kind == kPhaseOp) {
            g.c = d.c;
            g.s = d.s;
        } else {
            g.pivot = (uint32_t)__builtin_ctz(g.xc);
            g.c = d.pc.c;
            g.s = d.pc.s;
            g.base_imag = d.pc.base_imag;
            g.base_neg = d.pc.base_neg;
        }
        g.cf = (float)g.c;
        g.sf = (float)g.s;
        // bits outside the tile that every gate fixes to the same value
        if (i == 0) vcom = g.vo;
        mcom &= g.mo;
        mcom &= ~((g.vo ^ vcom) & mcom);
    }
    vcom &= mcom;
    const uint64_t fix = B.P | mcom;
    a.freemask = ((nq == 64 ? 0ull : (1ull << nq)) - 1ull) & ~fix;
    a.fixval = vcom;
    a.ntiles = 1ull << __builtin_popcountll(a.freemask);
    // the masks as runs; for the tile mask only the runs that the bits of a
    // thread's first element (below V kBlock) reach
    const uint32_t vec = c->f32 ? 4u : 2u;
    a.nfree_runs = mask_runs(a.freemask, a.free_runs, 64);
    if (a.nfree_runs && (a.free_runs[a.nfree_runs - 1] >> 16) + ((a.free_runs[a.nfree_runs - 1] >> 8) & 63u) !=
                            (uint32_t)__builtin_popcountll(a.freemask)) {
        set_last_error("tile enumeration mask has more runs than the pass kernel takes");
        return hipErrorInvalidValue;
    }
    a.ntile_runs = mask_runs(B.P, a.tile_runs, (uint32_t)__builtin_ctz(vec * (uint32_t)kBlock));
    const uint32_t cb = (uint32_t)__builtin_ctz(vec * (uint32_t)kBlock);   // tile bit of chunk index bit 0
    for (uint32_t j = 0; j < 3; ++j) a.kbit[j] = cb + j < t ? B.vec[cb + j] : 0ull;
    // A state that is one small tile (t <= kTinyBits) is one workgroup and the
    // pass is latency-bound: k_tile_tiny, one amplitude per thread (the steps
    // below are not used; its offsets are deposits of tmask: qubit tiles only).
    const bool tiny = t <= kTinyBits && a.ntiles == 1 && !B.skew;
    // register steps: runs of gates whose flipped tile bits fit in kSub bits
    a.nsteps = 0;
    uint32_t F = 0, g0 = 0;
    auto close = [&](uint32_t g1, bool lds) {
        auto& S = a.step[a.nsteps++];
        S.g0 = g0;
        S.g1 = g1;
        if (lds) {
            S.f[0] = S.f[1] = S.f[2] = kLdsStep;
            return;
        }
        // pad with tile bits that the step's gates select on, then high tile bits
        uint32_t want = 0;
        for (uint32_t k2 = g0; k2 < g1; ++k2) want |= a.g[k2].mc;
        for (uint32_t p = 0; p < t && __builtin_popcount(F) < (int)kSub; ++p)
            if (((want & ~F) >> p) & 1u) F |= 1u << p;
        for (int p = (int)t - 1; p >= 0 && __builtin_popcount(F) < (int)kSub; --p) F |= 1u << p;
        uint32_t qn = 0;
        for (uint32_t p = 0; p < t; ++p)
            if ((F >> p) & 1u) S.f[qn++] = p;
        for (uint32_t k2 = g0; k2 < g1; ++k2) local_tables(a.g[k2], S.f);
    };
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t xc = a.g[i].xc;
        // register steps assume that a one-bit partner mask is selected at 0
        const bool pivot_zero = __builtin_popcount(xc) != 1 || (((a.g[i].mc & xc) != 0) && ((a.g[i].vc & xc) == 0));
        // register arms for several flipped bits exist for SWAP and Pauli strings only
        const bool multi_ok = __builtin_popcount(xc) <= 1 || a.g[i].kind == (uint32_t)OP_SWAP ||
                              a.g[i].kind == (uint32_t)OP_PAULI;
        if ((uint32_t)__builtin_popcount(xc) > kSub || !pivot_zero || !multi_ok) {   // a step of its own, in LDS
            if (i > g0) close(i, false);
            g0 = i;
            close(i + 1, true);
            g0 = i + 1;
            F = 0;
            continue;
        }
        if ((uint32_t)__builtin_popcount(F | xc) > kSub) {
            close(i, false);
            g0 = i;
            F = 0;
        }
        F |= xc;
    }
    if (n > g0) close(n, false);
    const uint64_t g8 = ((a.ntiles + 7) / 8) * 8;
    const size_t es = c->f32 ? sizeof(float) : sizeof(double);
    const size_t lds = es * 2 * ((size_t)1 << t);
    ++c->passes;
    c->pass_gates += n;
    // algorithmic bytes: re and im of every tile amplitude read and written
    timing_begin(c, c->f32 ? "tile_pass32" : "tile_pass", 4.0 * (double)es * (double)((uint64_t)1 << t) * (double)a.ntiles);
    // groups per thread and dispatch: 2 when the tile has 2 kBlock groups or more
    const bool g2 = ((1u << t) / kLoc) >= 2 * kBlock;
    const dim3 grid((uint32_t)g8), blk(kBlock);
    const size_t tiny_lds = 2 * lds + sizeof(TileGate) * kMaxPass;
    if (c->f32) {
        if (tiny) hipLaunchKernelGGL(k_tile_tiny<float>, dim3(1), dim3(1u << t), tiny_lds, c->stream, c->f32_re, c->f32_im, a);
        else if (g2) hipLaunchKernelGGL((k_tile<float, 2>), grid, blk, lds, c->stream, c->f32_re, c->f32_im, a);
        else hipLaunchKernelGGL((k_tile<float, 1>), grid, blk, lds, c->stream, c->f32_re, c->f32_im, a);
    } else {
        if (tiny) hipLaunchKernelGGL(k_tile_tiny<double>, dim3(1), dim3(1u << t), tiny_lds, c->stream, c->d_re, c->d_im, a);
        else if (g2) hipLaunchKernelGGL((k_tile<double, 2>), grid, blk, lds, c->stream, c->d_re, c->d_im, a);
        else hipLaunchKernelGGL((k_tile<double, 1>), grid, blk, lds, c->stream, c->d_re, c->d_im, a);
    }
    timing_end(c);
    return hipGetLastError();
}

}  // namespace

namespace {

// Qubits a gate reads or writes: flipped, selected, and parity qubits.
uint64_t gate_support(const GateDesc& d) {
    const uint64_t par = d.kind == kPhaseOp ? d.zmask : d.pc.yz;
    const uint64_t flip = d.kind == kPhaseOp ? 0ull : d.X;
    return flip | d.M | par;
}
// Gates a, b commute when neither flips a qubit the other reads or writes:
// on their shared qubits both are then diagonal, so each is a sum over the
// shared qubits' values of projectors times operators on disjoint qubits.
bool gates_commute(const GateDesc& a, const GateDesc& b) {
    return (gate_flips(a) & gate_support(b)) == 0 && (gate_flips(b) & gate_support(a)) == 0;
}

// The order in which a packing gate queue is applied: passes packed greedily
// as flush() cuts them (qubits 0..f-1 plus flipped qubits, at most t; at most
// kMaxPass gates), a gate moving ahead only past gates it commutes with,
// looking `window` gates ahead.  The product of the gates is unchanged; the
// rounding is not: FP32 queues pack by default, FP64 contexts (whose default
// is the reference's gate order, bit for bit) only when asked to, for gate
// lists whose tolerance is a bound (amplitudes within 1e-12) rather than
// bit-identity.
void pack_gates(std::vector<GateDesc>& q, uint32_t t, uint32_t f, bool span, size_t window) {
    std::vector<GateDesc> rest(q.begin(), q.end()), out;
    out.reserve(q.size());
    size_t head = 0;   // rest[head..] not yet placed (erase from a vector front is linear; keep an index)
    std::vector<char> used(rest.size(), 0);
    while (head < rest.size()) {
        PassSpace ps;
        ps.low = (1ull << f) - 1ull;
        ps.span = span;
        uint32_t count = 0;
        bool took = true;
        while (took && count < kMaxPass && head < rest.size()) {
            took = false;
            std::vector<size_t> skipped;
            for (size_t j = head, seen = 0; j < rest.size() && seen < window; ++j) {
                if (used[j]) continue;
                ++seen;
                bool ok = true;
                for (size_t i = 0; ok && i < skipped.size(); ++i) ok = gates_commute(rest[skipped[i]], rest[j]);
                if (!ok || !ps.add(gate_flips(rest[j]), pass_sel(rest[j]), t - f)) {
                    skipped.push_back(j);
                    continue;
                }
                ++count;
                out.push_back(rest[j]);
                used[j] = 1;
                took = true;
                break;
            }
            while (head < rest.size() && used[head]) ++head;
        }
        if (count == 0) {   // the first gate alone
            out.push_back(rest[head]);
            used[head] = 1;
            while (head < rest.size() && used[head]) ++head;
        }
    }
    q.swap(out);
}

}  // namespace

namespace {

// The queue of a tiny_prog state as programs of up to kProgMax operations, in
// queue order.  A reference to an outcome that a measurement of the same
// program writes becomes that op's index (read from LDS); other references
// read the outcome slot (written by an earlier launch).
hipError_t run_tiny_programs(suqa_hip_ctx* c, const std::vector<GateDesc>& q) {
    const uint32_t N = 1u << c->nq;
    const uint32_t nthr = std::max<uint32_t>(N, (uint32_t)kBlock);
    const size_t lds = sizeof(double) * 5 * (size_t)N;
    size_t i = 0;
    while (i < q.size()) {
        ProgArgs a;
        std::memset(&a, 0, sizeof(a));
        a.re = c->d_re;
        a.im = c->d_im;
        a.slots = c->d_slots;
        a.nq = c->nq;
        uint32_t meas_op[kSlots];   // slot -> op index of this program's measurement writing it
        uint32_t n = 0;
        // the op index of this program's measurement writing slots ref .. ref + nb - 1, kNoRef when
        // none of them is written here; -2 (error) when only some are, or not by consecutive ops
        auto localize = [&](uint32_t ref, uint32_t nb, uint32_t* op0) -> int {
            uint32_t hits = 0;
            for (uint32_t b = 0; b < nb; ++b) {
                const uint32_t s = (ref + b) % kSlots;
                if (meas_op[s] != kNoRef) ++hits;
            }
            if (hits == 0) return 0;
            const uint32_t o = meas_op[ref % kSlots];
            if (hits != nb || o == kNoRef) return -2;
            for (uint32_t b = 0; b < nb; ++b)
                if (meas_op[(ref + b) % kSlots] != o + b) return -2;
            *op0 = o;
            return 1;
        };
        for (uint32_t s = 0; s < kSlots; ++s) meas_op[s] = kNoRef;
        uint64_t alg = 0;
        for (; i < q.size() && n < kProgMax; ++i) {
            const GateDesc& d = q[i];
            // a call's measurements stay in one program, so that an operation reading
            // their outcomes reads them all from this program or all from the slots
            if (d.kind == kMeasOp && n > 0 && d.meas_group <= kProgMax && n + d.meas_group > kProgMax) break;
            ProgOp& o = a.op[n];
            o.kf = (uint32_t)d.kind;
            if (d.kind == kMeasOp) {
                o.x = 1u << d.q;
                o.m = (uint32_t)d.M;
                o.yz = d.slot;
                o.c = d.c;
                if (d.reset) o.kf |= kPfReset;
                meas_op[d.slot % kSlots] = n;
                alg += 40ull * N;
            } else if (d.kind == kWOp) {
                o.x = d.wmask;
                o.m = d.omask;
                o.v = d.b0 | (d.ob0 << 8);
                o.yz = d.L;
                uint64_t tb = 0;
                std::memcpy(&tb, &d.tab, sizeof(d.tab));
                std::memcpy(&o.c, &tb, sizeof(o.c));
                if (d.eold_bits) {
                    uint32_t op0 = 0;
                    const int lz = localize(d.eold_ref, d.eold_bits, &op0);
                    if (lz < 0) {
                        set_last_error("apply_W reads outcomes that this program writes out of order");
                        return hipErrorInvalidValue;
                    }
                    o.ref = lz ? op0 : d.eold_ref;
                    if (lz) o.kf |= kPfLocal;
                    o.kf |= d.eold_bits << 16;
                } else {
                    o.ref = d.eold;
                }
                alg += 32ull * N;
            } else {
                o.m = (uint32_t)d.M;
                o.v = (uint32_t)d.V;
                if (d.kind == kPhaseOp) {
                    o.yz = (uint32_t)d.zmask;
                    o.c = d.c;
                    o.s = d.s;
                } else {
                    o.x = (uint32_t)d.X;
                    o.yz = (uint32_t)d.pc.yz;
                    o.c = d.pc.c;
                    o.s = d.pc.s;
                    o.kf |= (uint32_t)__builtin_ctz(o.x) << 9;
                    if (d.pc.base_imag) o.kf |= kPfImag;
                    if (d.pc.base_neg) o.kf |= kPfNeg;
                }
                if (d.cond_ref != kNoRef) {
                    uint32_t op0 = 0;
                    const int lz = localize(d.cond_ref, 1, &op0);
                    o.ref = lz ? op0 : d.cond_ref;
                    o.kf |= kPfCond | (lz ? kPfLocal : 0u);
                }
                alg += (uint64_t)d.alg_bytes;
            }
            ++n;
        }
        a.nops = n;
        ++c->passes;
        c->pass_gates += n;
        timing_begin(c, "tiny_prog", (double)alg);
        hipLaunchKernelGGL(k_tiny_prog, dim3(1), dim3(nthr), lds, c->stream, a);
        timing_end(c);
        if (hipError_t e = hipGetLastError()) return e;
    }
    return hipSuccess;
}

}  // namespace

void queue_measure(suqa_hip_ctx* c, const uint32_t* qs, const double* rds, uint32_t n, bool reset, uint32_t gc,
                   uint32_t slot) {
    for (uint32_t k = 0; k < n; ++k) {
        GateDesc d{};
        d.kind = kMeasOp;
        d.q = qs[k];
        d.c = rds[k];
        d.M = gc;
        d.reset = reset ? 1u : 0u;
        d.slot = (slot + k) % kSlots;
        d.meas_group = k == 0 ? n : 0u;
        d.tag = "measure";
        c->queue.push_back(d);
    }
}

// The queue in passes (programs for a tiny_prog state).  keep_tail, for the
// flush a long queue triggers: the gates of the queue's last passes stay
// queued, in their applied order, to join the operations queued next (packed:
// the passes formed from the last 2 kPackWindow gates, which had few gates to
// choose from; in list order: the last pass, which later gates may fill).
// Either way the gates apply in an order flush() could have produced.
constexpr size_t kPackWindow = 64;
hipError_t run_queue(suqa_hip_ctx* c, bool keep_tail) {
    if (c->queue.empty()) return hipSuccess;
    std::vector<GateDesc> q;
    q.swap(c->queue);
    if (tiny_prog(c)) return run_tiny_programs(c, q);
    const uint32_t t = std::min(c->tile_bits, c->nq);
    const uint32_t f = std::min(c->f32 ? kPassLowFixed32 : kLowFixed, t);   // whole 128-byte lines
    const bool span = c->span_tiles && c->nq > t;   // a state of one tile needs no span
    if (c->gate_reorder) pack_gates(q, t, f, span, kPackWindow);
    size_t i = 0;
    while (i < q.size()) {
        if (keep_tail && c->gate_reorder && q.size() - i < 2 * kPackWindow) break;
        PassSpace ps;
        ps.low = (1ull << f) - 1ull;
        ps.span = span;
        size_t j = i;
        while (j < q.size() && j - i < kMaxPass && ps.add(gate_flips(q[j]), pass_sel(q[j]), t - f)) ++j;
        if (keep_tail && j == q.size()) break;
        hipError_t e;
        if (c->f32) {   // float states have no one-gate sweeps: a lone gate is a pass of one
            if (j == i) {
                set_last_error("a gate's partner qubits do not fit the FP32 pass tile");
                return hipErrorInvalidValue;
            }
            e = run_pass(c, q.data() + i, (uint32_t)(j - i), t, f, span);
        } else if (j - i <= 1 || t < kSub) {   // a lone gate, one whose partner bits exceed the tile, or a tiny state
            e = run_single(c, q[i]);
            j = i + 1;
        } else {
            e = run_pass(c, q.data() + i, (uint32_t)(j - i), t, f, span);
        }
        if (e != hipSuccess) return e;
        i = j;
    }
    if (i < q.size()) c->queue.assign(q.begin() + (std::ptrdiff_t)i, q.end());
    return hipSuccess;
}

hipError_t flush(suqa_hip_ctx* c) { return run_queue(c, false); }

hipError_t flush_auto(suqa_hip_ctx* c) { return run_queue(c, true); }

}  // namespace sqe
