// kernels.hip -- FP64 state-vector sweeps for gfx950 (MI355X).
//
// Layout: SoA re[N], im[N] in one HBM allocation (complex_defines.cuh:22-45).
// Every kernel is a one-shot sweep over a selection of basis indices
// (engine.hpp: Sel): one thread per item, no grid-stride loop.  When bit 0 of
// the index is free an item is two consecutive amplitudes moved with 16-byte
// nontemporal loads and stores, so each wave instruction covers 1 KiB
// contiguous of one array.  Blocks are renumbered so that each of the 8 XCDs
// streams one contiguous eighth of the selection (block b runs on XCD b % 8).
// Measured on MI355X, this shape moves 6.7 TB/s in a plain copy, against
// 5.3 TB/s for grid-stride or unrolled variants.  Arithmetic keeps the reference's operation order and
// is compiled with -ffp-contract=off, so amplitudes match the reference CPU
// path bit for bit (suqa_cpu.hpp:49-66 forms x*c - y*s, y*c + x*s).
#include "engine.hpp"

namespace sqe {

namespace {

constexpr double kTwoSqInv = 0.7071067811865475;   // suqa.cuh:57

typedef double dv2 __attribute__((ext_vector_type(2)));

// Inserts a zero at each fixed position (ascending), then ORs their values.
__device__ __forceinline__ uint64_t expand(uint64_t k, const Fixed& fb, uint64_t fval) {
    for (uint32_t j = 0; j < fb.n; ++j) {
        const uint64_t lo = (1ull << fb.pos[j]) - 1ull;
        k = (k & lo) | ((k & ~lo) << 1);
    }
    return k | fval;
}

// Item index of this thread; the grid is a multiple of 8 blocks and block b
// takes chunk (b % 8) * (G / 8) + b / 8.
__device__ __forceinline__ uint64_t item_id() {
    const uint64_t G = gridDim.x, b = blockIdx.x;
    const uint64_t chunk = (b & 7ull) * (G >> 3) + (b >> 3);
    return chunk * kBlock + threadIdx.x;
}

template <int VEC>
struct Vec;
template <>
struct Vec<1> {
    double x[1];
    __device__ __forceinline__ static Vec load(const double* p) {
        Vec v;
        v.x[0] = __builtin_nontemporal_load(p);
        return v;
    }
    __device__ __forceinline__ void store(double* p) const { __builtin_nontemporal_store(x[0], p); }
};
template <>
struct Vec<2> {
    double x[2];
    __device__ __forceinline__ static Vec load(const double* p) {
        const dv2 t = __builtin_nontemporal_load(reinterpret_cast<const dv2*>(p));
        Vec v;
        v.x[0] = t.x;
        v.x[1] = t.y;
        return v;
    }
    __device__ __forceinline__ void store(double* p) const {
        dv2 t;
        t.x = x[0];
        t.y = x[1];
        __builtin_nontemporal_store(t, reinterpret_cast<dv2*>(p));
    }
};

struct PairArgs {
    Fixed fb;
    uint64_t fval;
    uint64_t nitems;   // items of VEC amplitudes
    uint64_t xmask;
    PauliCoef pc;
};

// psi'_a = c psi_a + zeta s psi_b, zeta in {1,-1,i,-i}; t = +-s carries the sign.
__device__ __forceinline__ void pauli_out(double ar, double ai, double br, double bi, double c, double t,
                                          bool imag, double& orr, double& oi) {
    if (imag) {   // zeta = +-i: re = ar*c - (+-s)*bi, im = ai*c + (+-s)*br
        orr = ar * c + bi * (-t);
        oi = ai * c + br * t;
    } else {      // zeta = +-1
        orr = ar * c + br * t;
        oi = ai * c + bi * t;
    }
}

template <int OP, int VEC>
__global__ __launch_bounds__(kBlock) void k_pair(double* __restrict__ re, double* __restrict__ im,
                                                 const PairArgs a) {
    const uint64_t it = item_id();
    if (it >= a.nitems) return;
    const uint64_t ia = expand(it * VEC, a.fb, a.fval);
    const uint64_t ib = ia ^ a.xmask;
    const Vec<VEC> ar = Vec<VEC>::load(re + ia), ai = Vec<VEC>::load(im + ia);
    const Vec<VEC> br = Vec<VEC>::load(re + ib), bi = Vec<VEC>::load(im + ib);
    Vec<VEC> oar, oai, obr, obi;
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
        const double xr = ar.x[e], xi = ai.x[e], yr = br.x[e], yi = bi.x[e];
        if (OP == OP_SWAP) {
            oar.x[e] = yr; oai.x[e] = yi; obr.x[e] = xr; obi.x[e] = xi;
        } else if (OP == OP_Y) {   // suqa_cpu.hpp:116-118
            oar.x[e] = yi; oai.x[e] = yr * -1.0; obr.x[e] = xi * -1.0; obi.x[e] = xr;
        } else if (OP == OP_H) {   // suqa_cpu.hpp:195-198
            oar.x[e] = kTwoSqInv * (xr + yr);
            oai.x[e] = kTwoSqInv * (xi + yi);
            obr.x[e] = kTwoSqInv * (xr - yr);
            obi.x[e] = kTwoSqInv * (xi - yi);
        } else {                    // OP_PAULI
            const uint32_t par = (uint32_t)__popcll((ia + (uint64_t)e) & a.pc.yz) & 1u;
            const uint32_t neg_a = a.pc.base_neg ^ par ^ a.pc.ny_odd;
            const uint32_t neg_b = a.pc.base_neg ^ par;
            const double ta = neg_a ? -a.pc.s : a.pc.s;
            const double tb = neg_b ? -a.pc.s : a.pc.s;
            pauli_out(xr, xi, yr, yi, a.pc.c, ta, a.pc.base_imag, oar.x[e], oai.x[e]);
            pauli_out(yr, yi, xr, xi, a.pc.c, tb, a.pc.base_imag, obr.x[e], obi.x[e]);
        }
    }
    oar.store(re + ia);
    oai.store(im + ia);
    obr.store(re + ib);
    obi.store(im + ib);
}

// ---------------------------------------------------------------------------
// Dense-low-bit mode.  A fixed bit or partner bit below kLowBits leaves every
// 128-byte line (16 amplitudes) partly selected; skipping the unselected
// amplitudes only makes the accesses strided.  Instead, sweep whole lines:
// lanes 8m..8m+7 hold one 16-amplitude group as 8 double2, each element tests
// its own membership (role a: (j & M) == V; role b: ((j ^ X) & M) == V),
// and the partner j ^ X is read from the lane (j ^ X) & 15 of the same wave
// (lane xor (X & 15) >> 1, element xor X & 1), or from the paired group
// g ^ (X & ~15) that the same thread loads.  Unselected elements are written
// back unchanged, so every store covers full lines.
constexpr uint64_t kLowMask = 15;

struct PairDLArgs {
    Fixed fb;          // iteration: high fixed bits, plus the pivot of X & ~15
    uint64_t fval;
    uint64_t nitems;   // double2 items
    uint64_t M, V;     // role-a predicate of the gate
    uint64_t X;        // partner mask
    uint64_t xhigh;    // X & ~15
    int lanex;         // (X & 15) >> 1
    int ex;            // X & 1
    PauliCoef pc;
};

__device__ __forceinline__ Vec<2> shfl_pick(const Vec<2>& v, int lanex, int ex) {
    Vec<2> r;
    double a = v.x[0], b = v.x[1];
    if (lanex) {
        a = __shfl_xor(a, lanex, 64);
        b = __shfl_xor(b, lanex, 64);
    }
    r.x[0] = ex ? b : a;
    r.x[1] = ex ? a : b;
    return r;
}

// Output of element j (value x) given its partner k = j ^ X (value y).
template <int OP>
__device__ __forceinline__ void dl_out(uint64_t j, const PairDLArgs& a, double xr, double xi, double yr, double yi,
                                       double& orr, double& oi) {
    const bool inA = (j & a.M) == a.V;
    const bool inB = ((j ^ a.X) & a.M) == a.V;
    orr = xr;
    oi = xi;
    if (!(inA || inB)) return;
    if (OP == OP_SWAP) {
        orr = yr;
        oi = yi;
    } else if (OP == OP_Y) {            // suqa_cpu.hpp:116-118
        if (inA) { orr = yi; oi = yr * -1.0; }
        else { orr = yi * -1.0; oi = yr; }
    } else if (OP == OP_H) {            // suqa_cpu.hpp:195-198
        if (inA) { orr = kTwoSqInv * (xr + yr); oi = kTwoSqInv * (xi + yi); }
        else { orr = kTwoSqInv * (yr - xr); oi = kTwoSqInv * (yi - xi); }
    } else {                            // OP_PAULI: psi_j' = c psi_j + base (-1)^{par(k&YZ)} s psi_k
        const uint32_t par = (uint32_t)__popcll((j ^ a.X) & a.pc.yz) & 1u;
        const double t = (a.pc.base_neg ^ par) ? -a.pc.s : a.pc.s;
        pauli_out(xr, xi, yr, yi, a.pc.c, t, a.pc.base_imag, orr, oi);
    }
}

template <int OP>
__global__ __launch_bounds__(kBlock) void k_pair_dl(double* __restrict__ re, double* __restrict__ im,
                                                    const PairDLArgs a) {
    const uint64_t it0 = item_id();
    const bool ok = it0 < a.nitems;
    const uint64_t p = expand((ok ? it0 : 0) * 2, a.fb, a.fval);
    const bool two = a.xhigh != 0;
    const uint64_t q = p ^ a.xhigh;
    const Vec<2> pr = Vec<2>::load(re + p), pi = Vec<2>::load(im + p);
    Vec<2> qr = pr, qi = pi;
    if (two) {
        qr = Vec<2>::load(re + q);
        qi = Vec<2>::load(im + q);
    }
    // partners of p's elements live in q's registers of lane ^ lanex
    const Vec<2> yr = shfl_pick(qr, a.lanex, a.ex), yi = shfl_pick(qi, a.lanex, a.ex);
    Vec<2> opr, opi;
#pragma unroll
    for (int e = 0; e < 2; ++e)
        dl_out<OP>(p + (uint64_t)e, a, pr.x[e], pi.x[e], yr.x[e], yi.x[e], opr.x[e], opi.x[e]);
    if (two) {
        const Vec<2> zr = shfl_pick(pr, a.lanex, a.ex), zi = shfl_pick(pi, a.lanex, a.ex);
        Vec<2> oqr, oqi;
#pragma unroll
        for (int e = 0; e < 2; ++e)
            dl_out<OP>(q + (uint64_t)e, a, qr.x[e], qi.x[e], zr.x[e], zi.x[e], oqr.x[e], oqi.x[e]);
        if (ok) {
            oqr.store(re + q);
            oqi.store(im + q);
        }
    }
    if (ok) {
        opr.store(re + p);
        opi.store(im + p);
    }
}

struct PhaseArgs {
    Fixed fb;
    uint64_t fval;
    uint64_t nitems;
    double c, s;
    uint64_t zmask;    // s changes sign where par(i & zmask) == 1
    uint64_t M, V;     // dense-low mode: element i is rotated iff (i & M) == V
};

// util_rotate2 (suqa_cpu.hpp:49-54): re' = re*c - im*s, im' = im*c + re*s.
template <int VEC, bool DL>
__global__ __launch_bounds__(kBlock) void k_phase(double* __restrict__ re, double* __restrict__ im,
                                                  const PhaseArgs a) {
    const uint64_t it = item_id();
    if (it >= a.nitems) return;
    const uint64_t ia = expand(it * VEC, a.fb, a.fval);
    const Vec<VEC> xr = Vec<VEC>::load(re + ia), xi = Vec<VEC>::load(im + ia);
    Vec<VEC> orr, oi;
#pragma unroll
    for (int e = 0; e < VEC; ++e) {
        const uint64_t j = ia + (uint64_t)e;
        const uint32_t par = (uint32_t)__popcll(j & a.zmask) & 1u;
        const double s = par ? -a.s : a.s;
        const double r = xr.x[e], m = xi.x[e];
        if (DL && (j & a.M) != a.V) {
            orr.x[e] = r;
            oi.x[e] = m;
            continue;
        }
        orr.x[e] = r * a.c - m * s;
        oi.x[e] = m * a.c + r * s;
    }
    orr.store(re + ia);
    oi.store(im + ia);
}

struct PhaseListArgs {
    Fixed fb;
    uint64_t fval;
    uint64_t nitems;
    uint32_t q0, size_mask;
    double c[128];
    double s[128];
};

__global__ __launch_bounds__(kBlock) void k_phase_list(double* __restrict__ re, double* __restrict__ im,
                                                       const PhaseListArgs a) {
    const uint64_t it = item_id();
    if (it >= a.nitems) return;
    const uint64_t i = expand(it, a.fb, a.fval);
    const uint32_t k = (uint32_t)(i >> a.q0) & a.size_mask;
    const double c = a.c[k], s = a.s[k];
    const double r = re[i], m = im[i];
    re[i] = r * c - m * s;
    im[i] = m * c + r * s;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

struct NormArgs {
    Fixed fb;
    uint64_t fval;
    uint64_t nitems;
    uint64_t split;    // 0: one sum; else sums for (i & split) == 0 and != 0
    uint64_t M, V;     // dense-low mode: element i counts iff (i & M) == V
};

template <int VEC, bool DL>
__global__ __launch_bounds__(kBlock) void k_norm(const double* __restrict__ re, const double* __restrict__ im,
                                                 const NormArgs a, double* __restrict__ partials) {
    __shared__ double lds[2][kBlock / 64];
    double acc0 = 0.0, acc1 = 0.0;
    const uint64_t it = item_id();
    if (it < a.nitems) {
        const uint64_t ia = expand(it * VEC, a.fb, a.fval);
        const Vec<VEC> xr = Vec<VEC>::load(re + ia), xi = Vec<VEC>::load(im + ia);
#pragma unroll
        for (int e = 0; e < VEC; ++e) {
            const uint64_t j = ia + (uint64_t)e;
            const double r = xr.x[e], m = xi.x[e];
            const double p = (DL && (j & a.M) != a.V) ? 0.0 : r * r + m * m;
            if (a.split && (j & a.split)) acc1 += p;
            else acc0 += p;
        }
    }
    acc0 = wave_sum(acc0);
    acc1 = wave_sum(acc1);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lds[0][w] = acc0;
        lds[1][w] = acc1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int k = 0; k < kBlock / 64; ++k) {
            s0 += lds[0][k];
            s1 += lds[1][k];
        }
        partials[2 * blockIdx.x] = s0;
        partials[2 * blockIdx.x + 1] = s1;
    }
}

// Fixed-order sum of pairs of partials: block k reduces pairs
// [k * chunk, (k+1) * chunk) into out[2k], out[2k+1].
__global__ __launch_bounds__(kBlock) void k_sum_pairs(const double* __restrict__ in, uint64_t npairs,
                                                      uint64_t chunk, double* __restrict__ out) {
    __shared__ double lds[2][kBlock / 64];
    const uint64_t lo = (uint64_t)blockIdx.x * chunk;
    const uint64_t hi = lo + chunk < npairs ? lo + chunk : npairs;
    double s0 = 0.0, s1 = 0.0;
    for (uint64_t k = lo + threadIdx.x; k < hi; k += kBlock) {
        s0 += in[2 * k];
        s1 += in[2 * k + 1];
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    if ((threadIdx.x & 63) == 0) {
        lds[0][threadIdx.x >> 6] = s0;
        lds[1][threadIdx.x >> 6] = s1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = 0.0, t1 = 0.0;
        for (int k = 0; k < kBlock / 64; ++k) {
            t0 += lds[0][k];
            t1 += lds[1][k];
        }
        out[2 * blockIdx.x] = t0;
        out[2 * blockIdx.x + 1] = t1;
    }
}

// measure_qbit collapse (suqa_cpu.hpp:544-549 then :36-39): the discarded half
// becomes +0.0, the kept half is multiplied by scale = 1./norm.
__global__ __launch_bounds__(kBlock) void k_collapse(double* __restrict__ re, double* __restrict__ im,
                                                     uint64_t nitems, uint64_t qbit, uint64_t keepbit,
                                                     double scale) {
    const uint64_t it = item_id();
    if (it >= nitems) return;
    const uint64_t i = it * 2;
    Vec<2> r, m;
    const bool k0 = (i & qbit) == keepbit;
    const bool k1 = ((i + 1) & qbit) == keepbit;
    if (k0 || k1) {
        r = Vec<2>::load(re + i);
        m = Vec<2>::load(im + i);
    } else {
        r.x[0] = r.x[1] = m.x[0] = m.x[1] = 0.0;
    }
    r.x[0] = k0 ? r.x[0] * scale : 0.0;
    m.x[0] = k0 ? m.x[0] * scale : 0.0;
    r.x[1] = k1 ? r.x[1] * scale : 0.0;
    m.x[1] = k1 ? m.x[1] * scale : 0.0;
    r.store(re + i);
    m.store(im + i);
}

__global__ __launch_bounds__(kBlock) void k_scale(double* __restrict__ data, uint64_t nitems, double f) {
    const uint64_t it = item_id();
    if (it >= nitems) return;
    Vec<2> v = Vec<2>::load(data + 2 * it);
    v.x[0] *= f;
    v.x[1] *= f;
    v.store(data + 2 * it);
}

__global__ __launch_bounds__(kBlock) void k_init(double* __restrict__ data, uint64_t nitems) {
    const uint64_t it = item_id();
    if (it >= nitems) return;
    Vec<2> v;
    v.x[0] = it == 0 ? 1.0 : 0.0;
    v.x[1] = 0.0;
    v.store(data + 2 * it);
}

// qms::apply_W, dense form of qms.cuh:286-307; row[2*E] = (fs1, fs2) for E_old fixed.
__global__ __launch_bounds__(kBlock) void k_qms_W(double* __restrict__ re, double* __restrict__ im,
                                                  uint64_t half, const double* __restrict__ row,
                                                  uint32_t wmask, uint32_t b0) {
    const uint64_t j = item_id();
    if (j >= half) return;
    const uint64_t i = j + half;
    const uint32_t En = ((uint32_t)i & wmask) >> b0;
    const double fs1 = row[2 * En], fs2 = row[2 * En + 1];
    double t = re[j];
    const double ri = re[i];
    re[j] = fs2 * t + fs1 * ri;
    re[i] = fs1 * t - fs2 * ri;
    t = im[j];
    const double mi = im[i];
    im[j] = fs2 * t + fs1 * mi;
    im[i] = fs1 * t - fs2 * mi;
}

__global__ __launch_bounds__(kBlock) void k_copy(dv2* __restrict__ dst, const dv2* __restrict__ src,
                                                 uint64_t n16) {
    const uint64_t it = item_id();
    if (it >= n16) return;
    __builtin_nontemporal_store(__builtin_nontemporal_load(src + it), dst + it);
}

// One thread per item, rounded up to a multiple of 8 blocks (item_id()).
inline uint32_t grid_for(uint64_t nitems) {
    uint64_t g = (nitems + kBlock - 1) / kBlock;
    g = (g + 7) & ~7ull;
    return (uint32_t)(g ? g : 8);
}

inline bool bit0_free(const Sel& s) { return s.fb.n == 0 || s.fb.pos[0] != 0; }

inline uint64_t sel_mask(const Sel& s) {
    uint64_t m = 0;
    for (uint32_t k = 0; k < s.fb.n; ++k) m |= 1ull << s.fb.pos[k];
    return m;
}

}  // namespace

bool make_sel(uint32_t nq, uint64_t M, uint64_t V, Sel* out) {
    const uint64_t all = (nq >= 64) ? ~0ull : ((1ull << nq) - 1ull);
    if ((V & ~M) != 0ull) return false;       // (i & M) == V has no solution
    if ((M & ~all) != 0ull) {
        // fixed bits beyond the state: a solution needs those bits of V to be 0
        if ((V & ~all) != 0ull) return false;
        M &= all;
    }
    out->fb.n = 0;
    for (uint32_t b = 0; b < nq; ++b)
        if ((M >> b) & 1ull) out->fb.pos[out->fb.n++] = b;
    out->fval = V & M;
    out->count = (1ull << nq) >> out->fb.n;
    return true;
}

hipError_t launch_pair(suqa_hip_ctx* c, int op, const Sel& sel, uint64_t xmask, const PauliCoef& pc,
                       const char* tag, double alg_bytes) {
    const uint64_t M = sel_mask(sel), V = sel.fval;
    if (((M | xmask) & kLowMask) != 0) {
        PairDLArgs a;
        a.M = M;
        a.V = V;
        a.X = xmask;
        a.xhigh = xmask & ~kLowMask;
        a.lanex = (int)((xmask & kLowMask) >> 1);
        a.ex = (int)(xmask & 1ull);
        a.pc = pc;
        const uint64_t Mi = (M & ~kLowMask & ~a.xhigh) | (a.xhigh & (~a.xhigh + 1ull));
        Sel it;
        make_sel(c->nq, Mi, V & Mi, &it);
        a.fb = it.fb;
        a.fval = it.fval;
        a.nitems = it.count / 2;
        const uint32_t g = grid_for(a.nitems);
        timing_begin(c, tag, alg_bytes);
#define SQE_PAIR_DL(OPV) \
    case OPV: hipLaunchKernelGGL((k_pair_dl<OPV>), dim3(g), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, a); break;
        switch (op) {
            SQE_PAIR_DL(OP_SWAP)
            SQE_PAIR_DL(OP_Y)
            SQE_PAIR_DL(OP_H)
            SQE_PAIR_DL(OP_PAULI)
        default: break;
        }
#undef SQE_PAIR_DL
        timing_end(c);
        return hipGetLastError();
    }
    PairArgs a;
    a.fb = sel.fb;
    a.fval = sel.fval;
    a.xmask = xmask;
    a.pc = pc;
    const bool v2 = bit0_free(sel) && sel.count >= 2 && (xmask & 1ull) == 0ull;
    a.nitems = v2 ? sel.count / 2 : sel.count;
    const uint32_t g = grid_for(a.nitems);
    timing_begin(c, tag, alg_bytes);
#define SQE_PAIR(OPV)                                                                             \
    case OPV:                                                                                     \
        if (v2) hipLaunchKernelGGL((k_pair<OPV, 2>), dim3(g), dim3(kBlock), 0, c->stream, c->d_re, \
                                   c->d_im, a);                                                   \
        else hipLaunchKernelGGL((k_pair<OPV, 1>), dim3(g), dim3(kBlock), 0, c->stream, c->d_re,    \
                                c->d_im, a);                                                      \
        break;
    switch (op) {
        SQE_PAIR(OP_SWAP)
        SQE_PAIR(OP_Y)
        SQE_PAIR(OP_H)
        SQE_PAIR(OP_PAULI)
    default: break;
    }
#undef SQE_PAIR
    timing_end(c);
    return hipGetLastError();
}

hipError_t launch_phase(suqa_hip_ctx* c, const Sel& sel, double cs, double sn, uint64_t zmask,
                        const char* tag, double alg_bytes) {
    PhaseArgs a;
    a.c = cs;
    a.s = sn;
    a.zmask = zmask;
    a.M = sel_mask(sel);
    a.V = sel.fval;
    const bool dl = (a.M & kLowMask) != 0;
    Sel it = sel;
    if (dl) make_sel(c->nq, a.M & ~kLowMask, a.V & ~kLowMask, &it);
    a.fb = it.fb;
    a.fval = it.fval;
    const bool v2 = it.count >= 2;
    a.nitems = v2 ? it.count / 2 : it.count;
    const uint32_t g = grid_for(a.nitems);
    timing_begin(c, tag, alg_bytes);
    if (dl && v2) hipLaunchKernelGGL((k_phase<2, true>), dim3(g), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, a);
    else if (dl) hipLaunchKernelGGL((k_phase<1, true>), dim3(g), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, a);
    else if (v2) hipLaunchKernelGGL((k_phase<2, false>), dim3(g), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, a);
    else hipLaunchKernelGGL((k_phase<1, false>), dim3(g), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, a);
    timing_end(c);
    return hipGetLastError();
}

hipError_t launch_phase_list(suqa_hip_ctx* c, const Sel& sel, uint32_t q0, uint32_t size_mask,
                             const double* cs, const double* sn) {
    PhaseListArgs a;
    a.fb = sel.fb;
    a.fval = sel.fval;
    a.nitems = sel.count;
    a.q0 = q0;
    a.size_mask = size_mask;
    for (uint32_t k = 0; k <= size_mask; ++k) {
        a.c[k] = cs[k];
        a.s[k] = sn[k];
    }
    const uint32_t g = grid_for(a.nitems);
    timing_begin(c, "phase_list", 32.0 * (double)sel.count);
    hipLaunchKernelGGL(k_phase_list, dim3(g), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, a);
    timing_end(c);
    return hipGetLastError();
}

uint64_t partial_pairs_capacity(uint64_t N) {
    // level-1 pairs (one per block of a VEC=1 sweep over N) + level-2 pairs
    const uint64_t p1 = grid_for(N);
    return p1 + (p1 + kPartChunk - 1) / kPartChunk + 8;
}

hipError_t launch_norm(suqa_hip_ctx* c, const Sel& sel, uint64_t split_bit, const char* tag,
                       double alg_bytes) {
    NormArgs a;
    a.split = split_bit;
    a.M = sel_mask(sel);
    a.V = sel.fval;
    const bool dl = (a.M & kLowMask) != 0;
    Sel it = sel;
    if (dl) make_sel(c->nq, a.M & ~kLowMask, a.V & ~kLowMask, &it);
    a.fb = it.fb;
    a.fval = it.fval;
    const bool v2 = it.count >= 2;
    a.nitems = v2 ? it.count / 2 : it.count;
    const uint32_t g = grid_for(a.nitems);
    timing_begin(c, tag, alg_bytes);
    if (dl && v2) hipLaunchKernelGGL((k_norm<2, true>), dim3(g), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, a, c->d_partials);
    else if (dl) hipLaunchKernelGGL((k_norm<1, true>), dim3(g), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, a, c->d_partials);
    else if (v2) hipLaunchKernelGGL((k_norm<2, false>), dim3(g), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, a, c->d_partials);
    else hipLaunchKernelGGL((k_norm<1, false>), dim3(g), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, a, c->d_partials);
    timing_end(c);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // deterministic tree: g pairs -> ceil(g / kPartChunk) pairs -> 1 pair
    const double* in = c->d_partials;
    uint64_t np = g;
    timing_begin(c, "norm_final", 16.0 * g);
    if (np > kPartChunk) {
        const uint64_t g2 = (np + kPartChunk - 1) / kPartChunk;
        double* out = c->d_partials + 2 * (uint64_t)grid_for(c->N);
        hipLaunchKernelGGL(k_sum_pairs, dim3((uint32_t)g2), dim3(kBlock), 0, c->stream, in, np,
                           (uint64_t)kPartChunk, out);
        in = out;
        np = g2;
    }
    hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(kBlock), 0, c->stream, in, np, np, c->d_result);
    timing_end(c);
    return hipGetLastError();
}

hipError_t launch_collapse(suqa_hip_ctx* c, uint64_t qbit, uint32_t keep, double scale) {
    const uint64_t nitems = c->N / 2;
    // the discarded half is written without being read: 24 B per amplitude
    timing_begin(c, "collapse", 24.0 * (double)c->N);
    hipLaunchKernelGGL(k_collapse, dim3(grid_for(nitems)), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, nitems,
                       qbit, keep ? qbit : 0ull, scale);
    timing_end(c);
    return hipGetLastError();
}

hipError_t launch_scale(suqa_hip_ctx* c, double f) {
    const uint64_t nitems = c->N;   // 2N doubles as N double2
    timing_begin(c, "scale", 32.0 * (double)c->N);
    hipLaunchKernelGGL(k_scale, dim3(grid_for(nitems)), dim3(kBlock), 0, c->stream, c->d_data, nitems, f);
    timing_end(c);
    return hipGetLastError();
}

hipError_t launch_init(suqa_hip_ctx* c) {
    const uint64_t nitems = c->N;
    timing_begin(c, "init", 16.0 * (double)c->N);
    hipLaunchKernelGGL(k_init, dim3(grid_for(nitems)), dim3(kBlock), 0, c->stream, c->d_data, nitems);
    timing_end(c);
    return hipGetLastError();
}

hipError_t launch_qms_W(suqa_hip_ctx* c, const double* d_row, uint32_t wmask, uint32_t b0) {
    const uint64_t half = c->N / 2;
    timing_begin(c, "qms_W", 32.0 * (double)c->N);
    hipLaunchKernelGGL(k_qms_W, dim3(grid_for(half)), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, half, d_row,
                       wmask, b0);
    timing_end(c);
    return hipGetLastError();
}

hipError_t launch_copy(suqa_hip_ctx* c, void* dst, const void* src, uint64_t nbytes) {
    const uint64_t n16 = nbytes / 16;
    timing_begin(c, "copy", 2.0 * (double)nbytes);
    hipLaunchKernelGGL(k_copy, dim3(grid_for(n16)), dim3(kBlock), 0, c->stream, reinterpret_cast<dv2*>(dst),
                       reinterpret_cast<const dv2*>(src), n16);
    timing_end(c);
    return hipGetLastError();
}

}  // namespace sqe
