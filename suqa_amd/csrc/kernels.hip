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
#include "lanes.hpp"
#include "sweep_math.hpp"

namespace sqe {

namespace {

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

__device__ __forceinline__ Vec<2> shx2(const Vec<2>& v, int m) {
    Vec<2> r;
    r.x[0] = __shfl_xor(v.x[0], m, 64);
    r.x[1] = __shfl_xor(v.x[1], m, 64);
    return r;
}

// Strided mode: a wave serves 16 double2 items; lanes 0-15 load re_a,
// 16-31 im_a, 32-47 re_b, 48-63 im_b (one 16-byte stream per lane, the
// layout that runs at copy speed), and the partner values come over
// lane ^ 32 (same component) and lane ^ 48 (other component).
template <int OP>
__global__ __launch_bounds__(kBlock) void k_pair(double* __restrict__ re, double* __restrict__ im,
                                                 const PairArgs a) {
    const uint64_t t = item_id();
    const int lane = threadIdx.x & 63;
    const uint64_t it = (t >> 6) * 16 + (uint64_t)(lane & 15);
    const bool ok = it < a.nitems;
    const bool b_side = lane >= 32, im_lane = (lane >> 4) & 1;
    const uint64_t ia = expand((ok ? it : 0) * 2, a.fb, a.fval);
    const uint64_t x = b_side ? ia ^ a.xmask : ia;
    double* base = im_lane ? im : re;
    const Vec<2> v = Vec<2>::load(base + x);
    const Vec<2> ps = shx2(v, 32);
    Vec<2> pc;
    if (OP == OP_Y || OP == OP_PAULI) pc = shx2(v, 48);
    else pc = ps;
    Vec<2> o;
#pragma unroll
    for (int e = 0; e < 2; ++e)
        o.x[e] = pair_out<OP>(!b_side, im_lane, v.x[e], ps.x[e], pc.x[e], (x + (uint64_t)e) ^ a.xmask, a.pc);
    if (ok) o.store(base + x);
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
    // element ex of the partner pair
    r.x[0] = ex ? b : a;
    r.x[1] = ex ? a : b;
    return r;
}

// Output of element j (value x) given its partner k = j ^ X (value y).
// Dense-low mode, stream-split: with a partner group (xhigh != 0) lanes
// 0-15 / 16-31 / 32-47 / 48-63 load re_P / im_P / re_Q / im_Q, 16 double2
// (two 16-amplitude groups) each; without one, lanes 0-31 / 32-63 load re / im
// of 32 double2.  Partner of element j: same component at lane ^ side ^ lanex,
// other component at lane ^ side ^ comp ^ lanex, element ^ ex.
template <int OP>
__global__ __launch_bounds__(kBlock) void k_pair_dl(double* __restrict__ re, double* __restrict__ im,
                                                    const PairDLArgs a) {
    const uint64_t t = item_id();
    const int lane = threadIdx.x & 63;
    const bool two = a.xhigh != 0;
    const int per = two ? 16 : 32;
    const int SIDE = two ? 32 : 0, COMP = two ? 16 : 32;
    const uint64_t it = (t >> 6) * (uint64_t)per + (uint64_t)(lane & (per - 1));
    const bool ok = it < a.nitems;
    const bool q_side = two && lane >= 32;
    const bool im_lane = (lane & COMP) != 0;
    const uint64_t p = expand((ok ? it : 0) * 2, a.fb, a.fval);
    const uint64_t x = q_side ? p ^ a.xhigh : p;
    double* base = im_lane ? im : re;
    const Vec<2> v = Vec<2>::load(base + x);
    const Vec<2> ps = shfl_pick(v, SIDE ^ a.lanex, a.ex);
    const Vec<2> pc = shfl_pick(v, SIDE ^ COMP ^ a.lanex, a.ex);
    Vec<2> o;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const uint64_t j = x + (uint64_t)e;
        const bool inA = (j & a.M) == a.V;
        const bool inB = ((j ^ a.X) & a.M) == a.V;
        o.x[e] = (inA || inB) ? pair_out<OP>(inA, im_lane, v.x[e], ps.x[e], pc.x[e], j ^ a.X, a.pc) : v.x[e];
    }
    if (ok) o.store(base + x);
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
// Stream-split: a wave serves 32 double2 items, lanes 0-31 load re and
// lanes 32-63 im; each lane gets the other component over lane ^ 32.
template <bool DL>
__global__ __launch_bounds__(kBlock) void k_phase(double* __restrict__ re, double* __restrict__ im,
                                                  const PhaseArgs a) {
    const uint64_t t = item_id();
    const int lane = threadIdx.x & 63;
    const uint64_t it = (t >> 6) * 32 + (uint64_t)(lane & 31);
    const bool ok = it < a.nitems;
    const bool im_lane = lane >= 32;
    const uint64_t ia = expand((ok ? it : 0) * 2, a.fb, a.fval);
    double* base = im_lane ? im : re;
    const Vec<2> v = Vec<2>::load(base + ia);
    const Vec<2> o = shx2(v, 32);
    Vec<2> out;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const uint64_t j = ia + (uint64_t)e;
        const uint32_t par = (uint32_t)__popcll(j & a.zmask) & 1u;
        const double s = par ? -a.s : a.s;
        out.x[e] = phase_out(im_lane, v.x[e], o.x[e], a.c, s);
        if (DL && (j & a.M) != a.V) out.x[e] = v.x[e];
    }
    if (ok) out.store(base + ia);
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

// the lane-xor butterfly over o = 32, 16, ..., 1 (every lane ends with the
// same sum), by DPP moves and permlane swaps instead of LDS permutes
__device__ __forceinline__ double wave_sum(double v) { return wave_sum64(v); }

struct NormArgs {
    Fixed fb;
    uint64_t fval;
    uint64_t nitems;
    uint64_t split;    // 0: one sum; else sums for (i & split) == 0 and != 0
    uint64_t M, V;     // dense-low mode: element i counts iff (i & M) == V
};

// Stream-split like k_phase: each lane squares its own component.
template <bool DL>
__global__ __launch_bounds__(kBlock) void k_norm(const double* __restrict__ re, const double* __restrict__ im,
                                                 const NormArgs a, double* __restrict__ partials) {
    __shared__ double lds[2][kBlock / 64];
    double acc0 = 0.0, acc1 = 0.0;
    const uint64_t t = item_id();
    const int lane = threadIdx.x & 63;
    const uint64_t it = (t >> 6) * 32 + (uint64_t)(lane & 31);
    if (it < a.nitems) {
        const uint64_t ia = expand(it * 2, a.fb, a.fval);
        const Vec<2> v = Vec<2>::load((lane >= 32 ? im : re) + ia);
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const uint64_t j = ia + (uint64_t)e;
            const double p = (DL && (j & a.M) != a.V) ? 0.0 : v.x[e] * v.x[e];
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

// Gates whose qubits lie in the active gc_mask.  The reference's dense loops
// (suqa_cpu.hpp:150-206, 268-317, 338-463) select i0 with
// (i0 & (gc | gate bits)) == gc and update the slots i0 | (subsets of the gate
// bits).  A gate bit inside gc is already set in i0, so slots coincide, and
// the loop body's reads and writes, done in order on the same addresses, make
// another operator than the gate's (H zeroes the selection, a swap becomes a
// controlled X, an XX rotation turns twice).  One thread per selected i0 runs
// the body of the reference's case on the slots k0 = i0, k1 = i0 | m1,
// k2 = i0 | m2, k3 = k1 | k2 through memory, in the reference's order.
__device__ __forceinline__ void rot2m(double* a, double* b, double c, double s) {   // suqa_cpu.hpp:49-54
    const double cpy = *a;
    *a = cpy * c - (*b) * s;
    *b = (*b) * c + cpy * s;
}
__device__ __forceinline__ void rot4m(double* a, double* b, double* c, double* d, double ct, double st) {
    rot2m(a, b, ct, st);
    rot2m(c, d, ct, st);
}
__device__ __forceinline__ void swapm(double* a, double* b) {
    const double cpy = *a;
    *a = *b;
    *b = cpy;
}

struct OverlapArgs {
    Fixed fb;
    uint64_t fval, nitems;
    uint64_t m1, m2;
    int op;              // kOvH, kOvSwap, kOvPauli1, kOvPauli2
    uint32_t p1, p2;     // Pauli types (sorted as the reference sorts them)
    double c, s;
};

__global__ __launch_bounds__(kBlock) void k_overlap(double* re, double* im, const OverlapArgs a) {
    const uint64_t it = item_id();
    if (it >= a.nitems) return;
    const uint64_t k0 = expand(it, a.fb, a.fval);
    const uint64_t k1 = k0 | a.m1, k2 = k0 | a.m2, k3 = k1 | k2;
    const double ct = a.c, st = a.s;
    switch (a.op) {
    case kOvH: {   // suqa_cpu.hpp:186-198
        const double a0r = re[k0], a0i = im[k0], a1r = re[k1], a1i = im[k1];
        re[k0] = kTwoSqInv * (a0r + a1r);
        im[k0] = kTwoSqInv * (a0i + a1i);
        re[k1] = kTwoSqInv * (a0r - a1r);
        im[k1] = kTwoSqInv * (a0i - a1i);
        break;
    }
    case kOvSwap:   // suqa_cpu.hpp:303-309
        swapm(&re[k1], &re[k2]);
        swapm(&im[k1], &im[k2]);
        break;
    case kOvPauli1:   // suqa_cpu.hpp:360-370
        if (a.p1 == 1) rot4m(&re[k0], &im[k1], &re[k1], &im[k0], ct, st);
        else if (a.p1 == 2) rot4m(&re[k0], &re[k1], &im[k0], &im[k1], ct, -st);
        else if (a.p1 == 3) rot4m(&re[k0], &im[k0], &im[k1], &re[k1], ct, st);
        break;
    default:   // kOvPauli2, suqa_cpu.hpp:422-440
        if (a.p1 == 1 && a.p2 == 1) {
            rot4m(&re[k0], &im[k3], &re[k3], &im[k0], ct, st);
            rot4m(&re[k1], &im[k2], &re[k2], &im[k1], ct, st);
        } else if (a.p1 == 2 && a.p2 == 2) {
            rot4m(&re[k0], &im[k3], &re[k3], &im[k0], ct, -st);
            rot4m(&re[k1], &im[k2], &re[k2], &im[k1], ct, st);
        } else if (a.p1 == 3 && a.p2 == 3) {
            rot4m(&re[k0], &im[k0], &re[k3], &im[k3], ct, st);
            rot4m(&re[k1], &im[k1], &re[k2], &im[k2], ct, -st);
        } else if (a.p1 == 1 && a.p2 == 2) {
            rot4m(&re[k0], &re[k3], &im[k0], &im[k3], ct, -st);
            rot4m(&re[k2], &re[k1], &im[k2], &im[k1], ct, -st);
        } else if (a.p1 == 1 && a.p2 == 3) {
            rot4m(&re[k0], &im[k1], &re[k1], &im[k0], ct, st);
            rot4m(&re[k2], &im[k3], &re[k3], &im[k2], ct, -st);
        } else if (a.p1 == 2 && a.p2 == 3) {
            rot4m(&re[k0], &re[k1], &im[k0], &im[k1], ct, -st);
            rot4m(&re[k2], &re[k3], &im[k2], &im[k3], ct, st);
        }
        break;
    }
}

// Checksum sums over the whole state, one amplitude per thread, into a
// (sum, sum) pair per block: mode 0 -> (sum re, sum im), mode 1 ->
// (sum i |psi_i|^2, 0).
__global__ __launch_bounds__(kBlock) void k_checksum(const double* __restrict__ re, const double* __restrict__ im,
                                                     uint64_t N, int mode, double* __restrict__ partials) {
    __shared__ double lds[2][kBlock / 64];
    double acc0 = 0.0, acc1 = 0.0;
    const uint64_t i = item_id();
    if (i < N) {
        const double r = re[i], m = im[i];
        if (mode == 0) {
            acc0 = r;
            acc1 = m;
        } else {
            acc0 = (double)i * (r * r + m * m);
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
    const uint64_t t = item_id();
    const int lane = threadIdx.x & 63;
    const uint64_t it = (t >> 6) * 32 + (uint64_t)(lane & 31);
    if (it >= nitems) return;
    const uint64_t i = it * 2;
    double* base = lane >= 32 ? im : re;
    const bool k0 = (i & qbit) == keepbit;
    const bool k1 = ((i + 1) & qbit) == keepbit;
    Vec<2> v;
    if (k0 || k1) v = Vec<2>::load(base + i);
    else v.x[0] = v.x[1] = 0.0;
    v.x[0] = k0 ? v.x[0] * scale : 0.0;
    v.x[1] = k1 ? v.x[1] * scale : 0.0;
    v.store(base + i);
}

// launch_measure_small (engine.hpp): one workgroup, the whole state held in
// registers, APT amplitudes per thread (amplitude y + T j in register j of
// thread y; T = blockDim.x = 256 for N <= 256, else min(N, 1024)).  The
// probabilities are a fixed-order block reduction: thread t < kBlock sums
// |psi_i|^2 for i = t, t + kBlock, t + 2 kBlock, ... in increasing i (the terms
// of up to 4 T consecutive indices pass through LDS at a time, one barrier
// each), then the wave butterfly and the four waves' sums in order; every
// thread reads the same partial sums, so the outcome and the scale are uniform.
struct MeasArgs {
    uint64_t N;
    uint32_t n, resets, gc, pad_;
    uint32_t q[kMeasMax];
    double rd[kMeasMax];
    uint32_t xq[kMeasMax];     // after op k, X on qubit xq[k] (over the gc bits) where xflags[k] reads 1;
    const uint32_t* xflags;    // xq[k] >= 32: none (the flips of another register, interleaved)
};

template <int APT>
__global__ __launch_bounds__(1024) void k_measure_small(double* __restrict__ re, double* __restrict__ im,
                                                        const MeasArgs a, uint32_t* __restrict__ out) {
    constexpr int kPB = APT < 4 ? APT : 4;   // registers whose terms share one pass through LDS
    __shared__ double pbuf[1024 * kPB];      // |psi|^2 of kPB T consecutive indices
    __shared__ double xr[1024], xi[1024];   // the reset's exchange
    __shared__ double red[2][2][kBlock / 64];
    const uint32_t T = blockDim.x, y = threadIdx.x;
    const uint32_t N = (uint32_t)a.N;
    const uint32_t tb = (uint32_t)__builtin_ctz(T);
    double vr[APT], vi[APT];
#pragma unroll
    for (int j = 0; j < APT; ++j) {
        const uint32_t i = y + T * (uint32_t)j;
        vr[j] = i < N ? re[i] : 0.0;
        vi[j] = i < N ? im[i] : 0.0;
    }
    // X on the qubit of bit qb over the indices with the other gc bits set
    // (uniform qb; every thread calls it: barriers inside)
    auto flip = [&](uint32_t qb) {
        const uint32_t g = a.gc & ~qb;   // both ends of a pair meet it or neither does
        if (qb >= T) {   // the partner in this thread's register j ^ (qb / T)
            const uint32_t h = qb >> tb;
#pragma unroll
            for (int hh = 1; hh < APT; hh <<= 1) {
                if (h != (uint32_t)hh) continue;
#pragma unroll
                for (int j = 0; j < APT; ++j) {
                    if (j & hh) continue;
                    if (((y + T * (uint32_t)j) & g) != g) continue;
                    const double tr = vr[j], ti = vi[j];
                    vr[j] = vr[j | hh];
                    vi[j] = vi[j | hh];
                    vr[j | hh] = tr;
                    vi[j | hh] = ti;
                }
            }
        } else if constexpr (APT <= 4) {   // the partner in thread y ^ qb, same register: every
            double pp[APT];                  // register of one component through pbuf at a time
#pragma unroll
            for (int j = 0; j < APT; ++j) pbuf[y + T * (uint32_t)j] = vr[j];
            __syncthreads();
#pragma unroll
            for (int j = 0; j < APT; ++j) pp[j] = pbuf[(y ^ qb) + T * (uint32_t)j];
            __syncthreads();
#pragma unroll
            for (int j = 0; j < APT; ++j)
                if (((y + T * (uint32_t)j) & g) == g) vr[j] = pp[j];
#pragma unroll
            for (int j = 0; j < APT; ++j) pbuf[y + T * (uint32_t)j] = vi[j];
            __syncthreads();
#pragma unroll
            for (int j = 0; j < APT; ++j) pp[j] = pbuf[(y ^ qb) + T * (uint32_t)j];
            __syncthreads();
#pragma unroll
            for (int j = 0; j < APT; ++j)
                if (((y + T * (uint32_t)j) & g) == g) vi[j] = pp[j];
        } else {         // the partner in thread y ^ qb, same register
#pragma unroll
            for (int j = 0; j < APT; ++j) {
                xr[y] = vr[j];
                xi[y] = vi[j];
                __syncthreads();
                const double pr = xr[y ^ qb], pi = xi[y ^ qb];
                __syncthreads();
                if (((y + T * (uint32_t)j) & g) == g) {
                    vr[j] = pr;
                    vi[j] = pi;
                }
            }
        }
    };
    uint32_t rb = 0;
    for (uint32_t k = 0; k < a.n; ++k) {
        const uint32_t q = a.q[k];
        const uint32_t qb = 1u << q;
        double s0 = 0.0, s1 = 0.0;
        if (N <= (uint32_t)kBlock) {   // one term per thread: thread y holds amplitude y
            const double p = vr[0] * vr[0] + vi[0] * vi[0];
            if (y < N) {
                if (y & qb) s1 += p;
                else s0 += p;
            }
        } else {
#pragma unroll
            for (int j0 = 0; j0 < APT; j0 += kPB) {
                if (j0) __syncthreads();   // the previous block's terms are summed
#pragma unroll
                for (int jj = 0; jj < kPB; ++jj) pbuf[y + T * (uint32_t)jj] = vr[j0 + jj] * vr[j0 + jj] + vi[j0 + jj] * vi[j0 + jj];
                __syncthreads();
                if (y < (uint32_t)kBlock)
                    for (uint32_t u = y; u < T * (uint32_t)kPB; u += kBlock) {   // index u + T j0, increasing
                        const double p = pbuf[u];
                        if ((u + T * (uint32_t)j0) & qb) s1 += p;
                        else s0 += p;
                    }
            }
        }
        if (y < (uint32_t)kBlock) {
            s0 = wave_sum(s0);
            s1 = wave_sum(s1);
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
        const uint32_t cbit = __builtin_amdgcn_readfirstlane(a.rd[k] < p1 ? 1u : 0u);   // suqa.cu:690
        const double scale = 1. / sqrt(cbit ? p1 : p0);  // set_ampl_to_zero + vnormalize (suqa.cu:694-695)
        const uint32_t keep = cbit ? qb : 0u;
#pragma unroll
        for (int j = 0; j < APT; ++j) {
            const bool kp = ((y + T * (uint32_t)j) & qb) == keep;
            vr[j] = kp ? vr[j] * scale : 0.0;
            vi[j] = kp ? vi[j] * scale : 0.0;
        }
        if (y == 0) out[k] = cbit;
        if (((a.resets >> k) & 1u) && cbit) flip(qb);   // apply_reset: X on q over the other gc bits (suqa.cu:735-737)
        if (a.xq[k] < 32u) {   // then the flip of another register's qubit by its outcome
            const uint32_t f = __builtin_amdgcn_readfirstlane(
                __hip_atomic_load(a.xflags + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            if (f) flip(1u << a.xq[k]);
        }
    }
#pragma unroll
    for (int j = 0; j < APT; ++j) {
        const uint32_t i = y + T * (uint32_t)j;
        if (i < N) {
            re[i] = vr[j];
            im[i] = vi[j];
        }
    }
}

// launch_x_if: X on qubit q[k] over the indices with the gc bits set where
// the outcome flags[k] reads 1 (the flips of a measured register), pairs
// (i, i | 2^q) with one thread each; the qubits in order, a barrier between
// them (one workgroup: small states only).
struct XifArgs {
    uint32_t n, gc;
    uint32_t q[kMeasMax];
    const uint32_t* flags;
};
__global__ __launch_bounds__(1024) void k_x_if(double* __restrict__ re, double* __restrict__ im, uint64_t N,
                                               const XifArgs a) {
    for (uint32_t k = 0; k < a.n; ++k) {
        const uint32_t f = __builtin_amdgcn_readfirstlane(
            __hip_atomic_load(a.flags + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (!f) continue;   // uniform
        const uint64_t qb = 1ull << a.q[k];
        const uint64_t g = (uint64_t)a.gc & ~qb;
        for (uint64_t p = threadIdx.x; p < N / 2; p += blockDim.x) {
            const uint64_t i = ((p >> a.q[k]) << (a.q[k] + 1)) | (p & (qb - 1ull));
            if ((i & g) != g) continue;
            const uint64_t j = i | qb;
            const double tr = re[i], ti = im[i];
            re[i] = re[j];
            im[i] = im[j];
            re[j] = tr;
            im[j] = ti;
        }
        __syncthreads();
    }
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
                                                  uint64_t half, const double* __restrict__ tab,
                                                  uint32_t wmask, uint32_t b0, uint32_t omask, uint32_t ob0,
                                                  uint32_t L) {
    const uint64_t j = item_id();
    if (j >= half) return;
    const uint64_t i = j + half;
    const uint32_t En = ((uint32_t)i & wmask) >> b0;
    const uint32_t Eo = ((uint32_t)i & omask) >> ob0;
    const double* row = tab + 2 * (uint64_t)Eo * L;
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
    if (c->tile_bits == 0) return launch_pair_now(c, op, sel, xmask, pc, tag, alg_bytes);
    GateDesc d{};
    d.kind = op;
    d.M = sel_mask(sel);
    d.V = sel.fval;
    d.X = xmask;
    d.pc = pc;
    d.tag = tag;
    d.alg_bytes = alg_bytes;
    c->queue.push_back(d);
    return c->queue.size() >= auto_flush_at(c) ? flush_auto(c) : hipSuccess;
}

hipError_t launch_phase(suqa_hip_ctx* c, const Sel& sel, double cs, double sn, uint64_t zmask,
                        const char* tag, double alg_bytes) {
    if (c->tile_bits == 0) return launch_phase_now(c, sel, cs, sn, zmask, tag, alg_bytes);
    GateDesc d{};
    d.kind = kPhaseOp;
    d.M = sel_mask(sel);
    d.V = sel.fval;
    d.c = cs;
    d.s = sn;
    d.zmask = zmask;
    d.tag = tag;
    d.alg_bytes = alg_bytes;
    c->queue.push_back(d);
    return c->queue.size() >= auto_flush_at(c) ? flush_auto(c) : hipSuccess;
}

hipError_t launch_pair_now(suqa_hip_ctx* c, int op, const Sel& sel, uint64_t xmask, const PauliCoef& pc,
                           const char* tag, double alg_bytes) {
    if (c->f32) return hipErrorInvalidValue;   // a gate queue over a float state: gates only
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
        const uint64_t per = a.xhigh ? 16 : 32;
        const uint32_t g = grid_for(((a.nitems + per - 1) / per) * 64);
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
    a.nitems = sel.count / 2;   // bits 0-3 are free here: double2 items
    const uint32_t g = grid_for(((a.nitems + 15) / 16) * 64);
    timing_begin(c, tag, alg_bytes);
#define SQE_PAIR(OPV) \
    case OPV: hipLaunchKernelGGL((k_pair<OPV>), dim3(g), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, a); break;
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

hipError_t launch_phase_now(suqa_hip_ctx* c, const Sel& sel, double cs, double sn, uint64_t zmask,
                            const char* tag, double alg_bytes) {
    if (c->f32) return hipErrorInvalidValue;   // a gate queue over a float state: gates only
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
    a.nitems = it.count / 2;   // N >= 2 and bit 0 free: double2 items
    const uint32_t g = grid_for(((a.nitems + 31) / 32) * 64);
    timing_begin(c, tag, alg_bytes);
    if (dl) hipLaunchKernelGGL((k_phase<true>), dim3(g), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, a);
    else hipLaunchKernelGGL((k_phase<false>), dim3(g), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, a);
    timing_end(c);
    return hipGetLastError();
}

hipError_t launch_phase_list(suqa_hip_ctx* c, const Sel& sel, uint32_t q0, uint32_t size_mask,
                             const double* cs, const double* sn) {
    if (c->f32) return hipErrorInvalidValue;   // a gate queue over a float state: gates only
    if (hipError_t e = flush(c)) return e;
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

hipError_t launch_norm_tree(const double* re, const double* im, uint32_t nq, const Sel& sel, uint64_t split_bit,
                            double* partials, double* result, hipStream_t st, suqa_hip_ctx* timed, const char* tag,
                            double alg_bytes) {
    NormArgs a;
    a.split = split_bit;
    a.M = sel_mask(sel);
    a.V = sel.fval;
    const bool dl = (a.M & kLowMask) != 0;
    Sel it = sel;
    if (dl) make_sel(nq, a.M & ~kLowMask, a.V & ~kLowMask, &it);
    a.fb = it.fb;
    a.fval = it.fval;
    a.nitems = it.count / 2;
    const uint32_t g = grid_for(((a.nitems + 31) / 32) * 64);
    if (timed) timing_begin(timed, tag, alg_bytes);
    if (dl) hipLaunchKernelGGL((k_norm<true>), dim3(g), dim3(kBlock), 0, st, re, im, a, partials);
    else hipLaunchKernelGGL((k_norm<false>), dim3(g), dim3(kBlock), 0, st, re, im, a, partials);
    if (timed) timing_end(timed);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // deterministic tree: g pairs -> ceil(g / kPartChunk) pairs -> 1 pair
    const double* in = partials;
    uint64_t np = g;
    if (timed) timing_begin(timed, "norm_final", 16.0 * g);
    if (np > kPartChunk) {
        const uint64_t g2 = (np + kPartChunk - 1) / kPartChunk;
        double* out = partials + 2 * (uint64_t)grid_for(1ull << nq);
        hipLaunchKernelGGL(k_sum_pairs, dim3((uint32_t)g2), dim3(kBlock), 0, st, in, np, (uint64_t)kPartChunk, out);
        in = out;
        np = g2;
    }
    hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(kBlock), 0, st, in, np, np, result);
    if (timed) timing_end(timed);
    return hipGetLastError();
}

hipError_t launch_norm(suqa_hip_ctx* c, const Sel& sel, uint64_t split_bit, const char* tag,
                       double alg_bytes) {
    if (c->f32) return hipErrorInvalidValue;   // a gate queue over a float state: gates only
    if (hipError_t e = flush(c)) return e;
    return launch_norm_tree(c->d_re, c->d_im, c->nq, sel, split_bit, c->d_partials, c->d_result, c->stream, c, tag,
                            alg_bytes);
}

hipError_t launch_overlap(suqa_hip_ctx* c, int op, const Sel& sel, uint64_t m1, uint64_t m2, uint32_t p1, uint32_t p2,
                          double cs, double sn) {
    if (c->f32) return hipErrorInvalidValue;   // a gate queue over a float state: gates only
    if (hipError_t e = flush(c)) return e;
    OverlapArgs a;
    a.fb = sel.fb;
    a.fval = sel.fval;
    a.nitems = sel.count;
    a.m1 = m1;
    a.m2 = m2;
    a.op = op;
    a.p1 = p1;
    a.p2 = p2;
    a.c = cs;
    a.s = sn;
    timing_begin(c, "overlap", 64.0 * (double)sel.count);
    hipLaunchKernelGGL(k_overlap, dim3(grid_for(a.nitems)), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, a);
    timing_end(c);
    return hipGetLastError();
}

hipError_t launch_checksum(suqa_hip_ctx* c) {
    if (c->f32) return hipErrorInvalidValue;   // a gate queue over a float state: gates only
    if (hipError_t e = flush(c)) return e;
    const uint32_t g = grid_for(c->N);
    for (int mode = 0; mode < 2; ++mode) {
        timing_begin(c, "checksum", 16.0 * (double)c->N);
        hipLaunchKernelGGL(k_checksum, dim3(g), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, c->N, mode,
                           c->d_partials);
        const double* in = c->d_partials;
        uint64_t np = g;
        if (np > kPartChunk) {
            const uint64_t g2 = (np + kPartChunk - 1) / kPartChunk;
            double* out = c->d_partials + 2 * (uint64_t)grid_for(c->N);
            hipLaunchKernelGGL(k_sum_pairs, dim3((uint32_t)g2), dim3(kBlock), 0, c->stream, in, np,
                               (uint64_t)kPartChunk, out);
            in = out;
            np = g2;
        }
        hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(kBlock), 0, c->stream, in, np, np, c->d_result + 2 * mode);
        timing_end(c);
        if (hipError_t e = hipGetLastError()) return e;
    }
    return hipSuccess;
}

hipError_t launch_collapse(suqa_hip_ctx* c, uint64_t qbit, uint32_t keep, double scale) {
    if (c->f32) return hipErrorInvalidValue;   // a gate queue over a float state: gates only
    if (hipError_t e = flush(c)) return e;
    const uint64_t nitems = c->N / 2;
    // the discarded half is written without being read: 24 B per amplitude
    timing_begin(c, "collapse", 24.0 * (double)c->N);
    hipLaunchKernelGGL(k_collapse, dim3(grid_for(((nitems + 31) / 32) * 64)), dim3(kBlock), 0, c->stream, c->d_re,
                       c->d_im, nitems, qbit, keep ? qbit : 0ull, scale);
    timing_end(c);
    return hipGetLastError();
}

hipError_t launch_measure_small(suqa_hip_ctx* c, const uint32_t* qs, const double* rds, uint32_t n, uint32_t resets,
                                uint32_t gc, uint32_t* out, const uint32_t* xqs, const uint32_t* xflags) {
    if (!out) out = c->d_flags;
    if (c->f32) return hipErrorInvalidValue;   // a gate queue over a float state: gates only
    if (hipError_t e = flush(c)) return e;
    if (n == 0 || n > kMeasMax || c->nq > kMeasSmallBits) return hipErrorInvalidValue;
    MeasArgs a{};
    a.N = c->N;
    a.n = n;
    a.resets = resets;
    a.gc = gc;
    a.xflags = xflags;
    for (uint32_t k = 0; k < n; ++k) {
        if (qs[k] >= c->nq || (xqs && xqs[k] >= c->nq)) return hipErrorInvalidValue;
        a.q[k] = qs[k];
        a.rd[k] = rds[k];
        a.xq[k] = xqs ? xqs[k] : ~0u;
    }
    timing_begin(c, "measure_small", 40.0 * (double)c->N * n);
    // the whole state in registers: 256 threads up to 256 amplitudes, then up to 1024 threads
    const uint32_t T = c->N <= (uint64_t)kBlock ? (uint32_t)kBlock : (uint32_t)std::min<uint64_t>(c->N, 1024);
    const uint32_t apt = (uint32_t)std::max<uint64_t>(1, c->N / T);
    switch (apt) {
    case 1: hipLaunchKernelGGL(k_measure_small<1>, dim3(1), dim3(T), 0, c->stream, c->d_re, c->d_im, a, out); break;
    case 2: hipLaunchKernelGGL(k_measure_small<2>, dim3(1), dim3(T), 0, c->stream, c->d_re, c->d_im, a, out); break;
    case 4: hipLaunchKernelGGL(k_measure_small<4>, dim3(1), dim3(T), 0, c->stream, c->d_re, c->d_im, a, out); break;
    case 8: hipLaunchKernelGGL(k_measure_small<8>, dim3(1), dim3(T), 0, c->stream, c->d_re, c->d_im, a, out); break;
    default: hipLaunchKernelGGL(k_measure_small<16>, dim3(1), dim3(T), 0, c->stream, c->d_re, c->d_im, a, out); break;
    }
    timing_end(c);
    return hipGetLastError();
}

hipError_t launch_x_if(suqa_hip_ctx* c, const uint32_t* qs, const uint32_t* flags, uint32_t n, uint32_t gc) {
    if (c->f32) return hipErrorInvalidValue;   // a gate queue over a float state: gates only
    if (hipError_t e = flush(c)) return e;
    if (n == 0) return hipSuccess;
    if (n > kMeasMax || c->nq > kMeasSmallBits) return hipErrorInvalidValue;
    XifArgs a{};
    a.n = n;
    a.gc = gc;
    a.flags = flags;
    for (uint32_t k = 0; k < n; ++k) {
        if (qs[k] >= c->nq) return hipErrorInvalidValue;
        a.q[k] = qs[k];
    }
    const uint32_t T = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(c->N / 2, (uint64_t)kBlock), 1024);
    timing_begin(c, "x_if", 32.0 * (double)c->N * n);
    hipLaunchKernelGGL(k_x_if, dim3(1), dim3(T), 0, c->stream, c->d_re, c->d_im, c->N, a);
    timing_end(c);
    return hipGetLastError();
}

hipError_t launch_scale(suqa_hip_ctx* c, double f) {
    if (c->f32) return hipErrorInvalidValue;   // a gate queue over a float state: gates only
    if (hipError_t e = flush(c)) return e;
    const uint64_t nitems = c->N;   // 2N doubles as N double2
    timing_begin(c, "scale", 32.0 * (double)c->N);
    hipLaunchKernelGGL(k_scale, dim3(grid_for(nitems)), dim3(kBlock), 0, c->stream, c->d_data, nitems, f);
    timing_end(c);
    return hipGetLastError();
}

hipError_t launch_init(suqa_hip_ctx* c) {
    if (c->f32) return hipErrorInvalidValue;   // a gate queue over a float state: gates only
    if (hipError_t e = flush(c)) return e;
    const uint64_t nitems = c->N;
    timing_begin(c, "init", 16.0 * (double)c->N);
    hipLaunchKernelGGL(k_init, dim3(grid_for(nitems)), dim3(kBlock), 0, c->stream, c->d_data, nitems);
    timing_end(c);
    return hipGetLastError();
}

hipError_t launch_qms_W(suqa_hip_ctx* c, const double* d_tab, uint32_t wmask, uint32_t b0, uint32_t omask,
                        uint32_t ob0, uint32_t L) {
    if (tiny_prog(c)) {   // queued beside the gates
        GateDesc d{};
        d.kind = kWOp;
        d.tab = d_tab;
        d.wmask = wmask;
        d.b0 = b0;
        d.omask = omask;
        d.ob0 = ob0;
        d.L = L;
        d.tag = "qms_W";
        c->queue.push_back(d);
        return c->queue.size() >= auto_flush_at(c) ? flush_auto(c) : hipSuccess;
    }
    if (hipError_t e = flush(c)) return e;
    const uint64_t half = c->N / 2;
    timing_begin(c, "qms_W", 32.0 * (double)c->N);
    hipLaunchKernelGGL(k_qms_W, dim3(grid_for(half)), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, half, d_tab,
                       wmask, b0, omask, ob0, L);
    timing_end(c);
    return hipGetLastError();
}

hipError_t launch_copy(suqa_hip_ctx* c, void* dst, const void* src, uint64_t nbytes) {
    if (c->f32) return hipErrorInvalidValue;   // a gate queue over a float state: gates only
    const uint64_t n16 = nbytes / 16;
    timing_begin(c, "copy", 2.0 * (double)nbytes);
    hipLaunchKernelGGL(k_copy, dim3(grid_for(n16)), dim3(kBlock), 0, c->stream, reinterpret_cast<dv2*>(dst),
                       reinterpret_cast<const dv2*>(src), n16);
    timing_end(c);
    return hipGetLastError();
}

}  // namespace sqe
