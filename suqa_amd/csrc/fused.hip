// fused.hip -- the precision pattern: FP32 state vector, gates fused into
// dense K-qubit blocks (K = 4..6, host-built in FP64, rounded to FP32) and
// applied with FP32 MFMA (v_mfma_f32_32x32x2f32) with FP32 accumulation.
// Norms, measurement probabilities, the collapse scale and the fidelity
// against an FP64 state are accumulated in FP64.
//
// Block passes.  The op list is cut, in order, into passes of consecutive ops
// whose target qubits fit one tile of t = 12 qubits (controls outside it);
// each pass is fused into K-qubit blocks inside its tile (fuse.cpp) and one
// kernel applies all of them: a workgroup loads a tile (2^12 amplitudes, the
// lowest 5 non-control qubits always included so a tile is whole 128-byte
// runs) into LDS, applies the blocks one after the other, and writes it back.
// A block on K tile qubits is Y = A X with X = [Xr; Xi] (2^(K+1) x 2^(12-K),
// the other tile bits as columns) and the real form A = [[Mr, -Mi], [Mi, Mr]];
// a wave holds one 32-row block of A in VGPRs (host-permuted to the MFMA lane
// order) and takes up to 2 of the (row block, 32-column block) pairs.  A block
// whose ops share a control applies only where the control is 1; a control
// common to the whole launch is fixed in the tile enumeration.
// MFMA 32x32x2 f32 operand layout: A[i][k] in lane i + 32k, B[k][j] in lane
// j + 32k, D[i][j] in lane j + 32 * ((i / 4) % 2), register 4 * (i / 8) + i % 4.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "deposit.hpp"
#include "engine.hpp"
#include "fused.hpp"

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double dv2f __attribute__((ext_vector_type(2)));

struct suqa_hip_f32 {
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t nq = 0;
    uint64_t N = 0;
    float* d_data = nullptr;   // re[N], im[N]
    float* d_re = nullptr;
    float* d_im = nullptr;
    double* d_partials = nullptr;
    uint64_t partial_cap = 0;  // pairs
    double* d_result = nullptr;
    double* h_result = nullptr;
    float* d_mats = nullptr;   // block matrices in MFMA lane order
    size_t mats_cap = 0;       // floats
    uint32_t passes = 0;       // 0: one block per launch, 1: block passes, 2: passes with sparse Pauli rotations,
                               // 3: gate passes (gates queued as on the FP64 engine, float tile passes)
    suqa_hip_ctx* gq = nullptr;   // mode 3: the gate queue over this state (sqe::f32_gate_queue)
    uint32_t gate_tile_bits = 12; // mode 3: tile qubits of a pass (12: 32 KiB of LDS, 13: 64 KiB)
    // per-launch HIP events of the block kernels (suqa_hip_f32_timing_*)
    bool timing = false;
    struct Rec {
        int stat;
        hipEvent_t e0, e1;
    };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> evpool;
    struct Stat {
        std::string name;
        uint64_t launches = 0;
        double ms = 0.0, bytes = 0.0, flops = 0.0;
    };
    std::vector<Stat> stats;
};

namespace sqf {
namespace {

thread_local std::string g_err;
int fail(int code, const std::string& m) {
    g_err = m;
    return code;
}
#define SQF_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess) return fail(SUQA_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int kBlock = 256;

// ---- per-launch timing ------------------------------------------------------
hipEvent_t take_event(suqa_hip_f32* c) {
    if (!c->evpool.empty()) {
        hipEvent_t e = c->evpool.back();
        c->evpool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}
// Brackets one launch: begin before, end after; bytes are the algorithmic HBM
// bytes of the launch, flops its MFMA flops.
int timing_begin(suqa_hip_f32* c, const std::string& name, double bytes, double flops) {
    if (!c->timing) return -1;
    size_t k = 0;
    while (k < c->stats.size() && c->stats[k].name != name) ++k;
    if (k == c->stats.size()) {
        c->stats.emplace_back();
        c->stats.back().name = name;
    }
    c->stats[k].launches++;
    c->stats[k].bytes += bytes;
    c->stats[k].flops += flops;
    suqa_hip_f32::Rec r{(int)k, take_event(c), take_event(c)};
    (void)hipEventRecord(r.e0, c->stream);
    c->recs.push_back(r);
    return (int)k;
}
void timing_end(suqa_hip_f32* c, int k) {
    if (k >= 0) (void)hipEventRecord(c->recs.back().e1, c->stream);
}
// After the stream has drained: the recorded durations into the totals.
void timing_collect(suqa_hip_f32* c) {
    for (auto& r : c->recs) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.e0, r.e1) == hipSuccess) c->stats[r.stat].ms += ms;
        c->evpool.push_back(r.e0);
        c->evpool.push_back(r.e1);
    }
    c->recs.clear();
}

__device__ __forceinline__ uint64_t item_id() {   // XCD-contiguous block remap (grid % 8 == 0)
    const uint64_t G = gridDim.x, b = blockIdx.x;
    return ((b & 7ull) * (G >> 3) + (b >> 3)) * kBlock + threadIdx.x;
}
inline uint32_t grid_for(uint64_t threads) {
    uint64_t g = (threads + kBlock - 1) / kBlock;
    g = (g + 7) & ~7ull;
    return (uint32_t)(g ? g : 8);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ---------------------------------------------------------------------------
// One block per launch: a workgroup takes one tile = the K block qubits plus the
// lowest other qubits, tile_bits1(K) in all (runs of at least 256 bytes),
// applies the block by MFMA and writes the tile back; the blocks come from
// fusing the whole op list (fuse.cpp).  11-qubit tiles (16 KiB) for K = 4 keep
// its registers low enough for 8 workgroups per CU; K = 5, 6 take 12-qubit
// tiles (32 KiB), which give them 2 (row block, column block) items per wave
// (measured: K = 6 from 1.33 to 1.03 ms per gate, K = 4 slower at 12).
constexpr uint32_t tile_bits1(uint32_t K) { return K == 4 ? 11u : 12u; }
constexpr uint32_t kItems = 2;        // (row block, column block) pairs per wave, at most
struct BlockArgs {
    uint32_t ntile;                    // K + L
    uint32_t L;                        // non-block bits per tile (columns = 2^L)
    uint32_t rowoff[1 << kMaxK];       // compact offset of block index r
    uint32_t coloff[128];              // compact offset of column c (at most 2^(tile_bits1(K) - K))
    uint64_t tmask;                    // the tile's state bits (compact order = ascending)
    uint64_t freemask;                 // state bits enumerated by the tile number
    uint64_t fixval;                   // the controls (all 1)
    uint64_t ntiles;
    const float* A;                    // [RB][D/2][64] MFMA-ordered stacked matrix
};

template <int K>
__global__ __launch_bounds__(kBlock) void k_block(float* __restrict__ re, float* __restrict__ im, const BlockArgs a) {
    constexpr int D = 2 << K;          // stacked dimension
    constexpr int RB = D / 32;         // 32-row blocks of A (1, 2, 4)
    constexpr int KS = D / 2;          // k-steps of 2
    extern __shared__ float lds[];
    const uint32_t tsz = 1u << a.ntile;
    float* lre = lds;
    float* lim = lds + tsz;
    const uint64_t tile = ((uint64_t)(blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3));
    if (tile >= a.ntiles) return;      // whole workgroup leaves together
    // A controlled block only visits tiles where every control is 1: the
    // others are left untouched.
    const uint64_t base = deposit(tile, a.freemask) | a.fixval;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;

    // 1. tile -> LDS (compact order), 16-byte chunks: chunk i of a thread is
    //    element 4 tid + 1024 i; every load is issued before the LDS writes
    constexpr uint32_t kCh = (1u << tile_bits1(K)) / 4 / kBlock;   // chunks per thread at full size
    constexpr uint32_t kIt = 1u << (tile_bits1(K) - 11);           // items per wave at full size
    const uint64_t toff = base | deposit(4u * tid, a.tmask);
    uint64_t ibit[3];   // state offsets of tile bits 10, 11, 12
#pragma unroll
    for (uint32_t j = 0; j < 3; ++j) ibit[j] = deposit(4u * kBlock << j, a.tmask);
    auto chunk_off = [&](uint32_t i) {
        return ((i & 1u) ? ibit[0] : 0ull) | ((i & 2u) ? ibit[1] : 0ull) | ((i & 4u) ? ibit[2] : 0ull);
    };
    constexpr uint32_t kBatch = kCh < 4 ? kCh : 4;   // loads in flight per thread
#pragma unroll
    for (uint32_t i0 = 0; i0 < kCh; i0 += kBatch) {
        floatx4 vr[kBatch], vi[kBatch];
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j)
            if (tid + (i0 + j) * kBlock < tsz / 4) {
                const uint64_t g = toff | chunk_off(i0 + j);
                vr[j] = __builtin_nontemporal_load(reinterpret_cast<const floatx4*>(re + g));
                vi[j] = __builtin_nontemporal_load(reinterpret_cast<const floatx4*>(im + g));
            }
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j)
            if (tid + (i0 + j) * kBlock < tsz / 4) {
                *reinterpret_cast<floatx4*>(lre + 4 * (tid + (i0 + j) * kBlock)) = vr[j];
                *reinterpret_cast<floatx4*>(lim + 4 * (tid + (i0 + j) * kBlock)) = vi[j];
            }
    }
    __syncthreads();

    // 2. MFMA: wave w owns row block rb and column blocks cb0 .. cb0+RB-1
    // (row block, 32-column block) items w + 4 it: 4 % RB == 0, so every item
    // of a wave has row block w % RB; at full size there are 8 items, 2 per wave
    const int rb = w % RB;
    const uint32_t items = (uint32_t)RB << (a.L - 5);
    float areg[KS];
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) areg[kk] = a.A[((size_t)rb * KS + kk) * 64 + lane];
    floatx16 acc[kIt];
    uint32_t coff[kIt];
    bool live[kIt];
#pragma unroll
    for (uint32_t it = 0; it < kIt; ++it) {
        const uint32_t item = (uint32_t)w + 4u * it;
        live[it] = item < items;
        // a wave without this item reads a valid column of its row block and stores nothing
        coff[it] = a.coloff[(live[it] ? item / RB : 0u) * 32u + (uint32_t)(lane & 31)];
        acc[it] = floatx16{0};
    }
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) {
        const int R = 2 * kk + (lane >> 5);
        const float* src = R < (1 << K) ? lre + a.rowoff[R] : lim + a.rowoff[R - (1 << K)];
        float b[kIt];
#pragma unroll
        for (uint32_t it = 0; it < kIt; ++it) b[it] = src[coff[it]];
#pragma unroll
        for (uint32_t it = 0; it < kIt; ++it)
            acc[it] = __builtin_amdgcn_mfma_f32_32x32x2f32(areg[kk], b[it], acc[it], 0, 0, 0);
    }
    __syncthreads();
#pragma unroll
    for (uint32_t it = 0; it < kIt; ++it) {
        if (!live[it]) continue;   // uniform
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int i = rb * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            if (i < (1 << K)) lre[a.rowoff[i] + coff[it]] = acc[it][v];
            else lim[a.rowoff[i - (1 << K)] + coff[it]] = acc[it][v];
        }
    }
    __syncthreads();

    // 3. LDS -> tile
#pragma unroll
    for (uint32_t i = 0; i < kCh; ++i)
        if (tid + i * kBlock < tsz / 4) {
            const uint64_t g = toff | chunk_off(i);
            const uint32_t e = 4 * (tid + i * kBlock);
            __builtin_nontemporal_store(*reinterpret_cast<const floatx4*>(lre + e), reinterpret_cast<floatx4*>(re + g));
            __builtin_nontemporal_store(*reinterpret_cast<const floatx4*>(lim + e), reinterpret_cast<floatx4*>(im + g));
        }
}

// ---------------------------------------------------------------------------
// Block passes: several fused blocks per HBM round trip.  A workgroup loads a
// tile of 2^t amplitudes (the t tile qubits, which hold every block qubit of
// the pass) into LDS, applies the pass's blocks one after the other by MFMA
// on the LDS-resident tile, and writes it back.  A block on K tile qubits
// multiplies 2^(t-K) columns (the other tile bits); the 4 waves share its
// 2^(K+1)/32 row blocks x 2^(t-K)/32 column blocks, at most kItems each.
constexpr uint32_t kPassTileBits = 12;   // 32 KiB of FP32 complex amplitudes per tile
constexpr uint32_t kPassBlocks = 16;     // blocks per launch

// LDS layout of a tile: amplitude e sits at lds_sw(e) = e ^ h(e), h the XOR of
// kSwVec[p] over the set bits p >= 5 of e.  h < 32 depends only on bits the
// XOR leaves alone, so lds_sw is a bijection of [0, 2^t) that keeps aligned
// groups of 4 together, and it is linear over GF(2).  Tile position p drives
// the LDS banks (address mod 32) by the vector kSwVec[p]: 32 lanes that span 5
// positions of independent vectors hit 32 distinct banks.  The vectors of
// positions 5..11 come from a search over every 4-, 5- and 6-position block of
// a 12-position tile: the 5 free positions picked by rank (see add_launch)
// always span a space of rank >= 4, at most a 2-way conflict (1.06 on average).
__host__ __device__ constexpr uint32_t kSwVecAt(uint32_t p) {
    constexpr uint32_t v[12] = {1, 2, 4, 8, 16, 26, 7, 22, 21, 3, 28, 13};
    return v[p];
}
__host__ __device__ inline uint32_t lds_sw(uint32_t e) {
    uint32_t h = 0;
#pragma unroll
    for (uint32_t p = 5; p < kPassTileBits; ++p)
        if ((e >> p) & 1u) h ^= kSwVecAt(p);
    return e ^ h;
}

struct PassBlock {
    uint32_t K;                               // 0: a swap along tile bit pbit, or a rotation (rot)
    uint32_t pbit;
    uint32_t rot;                             // 1: sparse Pauli rotation, pbit its lowest flipped bit (0: diagonal)
    uint32_t rflip, rpar;                     // rotation: flipped and parity tile bits
    float rc, rkr, rki;                       // rotation: psi' = rc psi(y) + rk (-1)^{|(y^rflip)&rpar|} psi(y^rflip)
    uint32_t rowsw[kMaxK];                    // lds_sw(1 << p) of the block positions, ascending
    uint32_t colsw[kPassTileBits - kMinK];    // lds_sw(1 << p) of column bit j (bits 0..4: the lanes)
    uint32_t pad_;
    uint64_t mo, vo;       // controls (outside the tile): applied where (base & mo) == vo
    uint64_t zo;           // rotation: parity qubits outside the tile (constant over a tile)
    const float* A;        // MFMA-ordered stacked matrix
};

// A per-tile scalar of a pass: multiplies the tiles whose base meets
// (base & mo) == vo (qubits outside the tile).
struct PassScal {
    uint64_t mo, vo;
    double re, im;
};
constexpr uint32_t kPassScal = 24;   // scalars per pass

struct PassArgs {
    uint32_t t, nblocks;
    uint64_t tmask;        // the tile's state bits (tile element e -> deposit(e, tmask))
    uint64_t freemask;     // state bits enumerated by the tile number
    uint64_t fixval;       // controls common to every block of the launch
    uint64_t ntiles;
    uint32_t K, nscal;     // launch width; scalars applied when the tile is stored
    PassScal sc[kPassScal];
    PassBlock b[kPassBlocks];
};

// v[k ^ p] in slot k (an involution: also its own inverse).
__device__ __forceinline__ floatx4 perm4(floatx4 v, uint32_t p) {
    const bool s1 = p & 1u, s2 = p & 2u;
    const float a0 = s1 ? v.y : v.x, a1 = s1 ? v.x : v.y, a2 = s1 ? v.w : v.z, a3 = s1 ? v.z : v.w;
    floatx4 r;
    r.x = s2 ? a2 : a0;
    r.y = s2 ? a3 : a1;
    r.z = s2 ? a0 : a2;
    r.w = s2 ? a1 : a3;
    return r;
}

template <int K>
__device__ __forceinline__ void pass_block(float* lre, float* lim, uint32_t t, const PassBlock& B) {
    constexpr int Dc = 1 << K;
    constexpr int RB = (2 * Dc) / 32;   // 32-row blocks of the stacked matrix (1, 2, 4)
    constexpr int KS = Dc;              // k-steps of 2 rows
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t nc = t - K;                   // column bits (>= 5)
    const uint32_t items = (uint32_t)RB << (nc - 5);
    // swizzled tile offset of block row r: an XOR of the row bits' swizzles (uniform)
    auto rows = [&](uint32_t r) {
        uint32_t o = 0;
#pragma unroll
        for (int q = 0; q < K; ++q)
            if ((r >> q) & 1u) o ^= B.rowsw[q];
        return o;
    };
    const int rb = w % RB;              // 4 % RB == 0: every item of a wave has this row block
    float areg[KS];
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) areg[kk] = B.A[((size_t)rb * KS + kk) * 64 + lane];
    uint32_t lanec = 0;                 // lane bits 0..4 -> column bits 0..4
#pragma unroll
    for (int j = 0; j < 5; ++j)
        if ((lane >> j) & 1) lanec ^= B.colsw[j];
    const uint32_t rd_half = (lane >> 5) ? B.rowsw[0] : 0u;   // input rows 2kk, 2kk+1: block bit 0
    const uint32_t wr_half = (lane >> 5) ? B.rowsw[2] : 0u;   // output rows i, i+4: block bit 2
    uint32_t cs[kItems];
    bool live[kItems];
#pragma unroll
    for (uint32_t it = 0; it < kItems; ++it) {
        const uint32_t item = (uint32_t)w + 4u * it;
        live[it] = item < items;
        // a wave without this item reads a valid column of its row block and stores nothing
        const uint32_t cb = live[it] ? item / RB : 0u;
        uint32_t c = lanec;
        for (uint32_t j = 5; j < nc; ++j)
            if ((cb >> (j - 5)) & 1u) c ^= B.colsw[j];
        cs[it] = c;
    }
    floatx16 acc[kItems];
#pragma unroll
    for (uint32_t it = 0; it < kItems; ++it) acc[it] = floatx16{0};
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) {
        // rows 2kk, 2kk+1 of the stacked matrix: both real (2kk < Dc) or both imaginary;
        // R0 is even, row R0 + 1 adds block bit 0 (the lane half)
        const uint32_t ro = rows((uint32_t)((2 * kk) % Dc)) ^ rd_half;
        float* src = (2 * kk < Dc) ? lre : lim;
        float b[kItems];
#pragma unroll
        for (uint32_t it = 0; it < kItems; ++it) b[it] = src[ro ^ cs[it]];
#pragma unroll
        for (uint32_t it = 0; it < kItems; ++it)
            acc[it] = __builtin_amdgcn_mfma_f32_32x32x2f32(areg[kk], b[it], acc[it], 0, 0, 0);
    }
    __syncthreads();   // every read of the block's input is done
#pragma unroll
    for (uint32_t it = 0; it < kItems; ++it) {
        if (!live[it]) continue;   // uniform
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            // output row rb*32 + 8(v/4) + 4(lane/32) + v%4: uniform up to the lane half
            const int i0 = rb * 32 + 8 * (v >> 2) + (v & 3);
            const bool imag = i0 >= Dc;   // rows i0 and i0 + 4 lie in the same half (Dc >= 16)
            const int r0 = imag ? i0 - Dc : i0;   // r0 % 8 < 4: row r0 + 4 adds block bit 2
            const uint32_t o = rows((uint32_t)r0) ^ wr_half ^ cs[it];
            if (imag) lim[o] = acc[it][v];
            else lre[o] = acc[it][v];
        }
    }
    __syncthreads();
}

// A controlled X on tile position p (rowsw[0] = lds_sw(1 << p)): swaps the
// two halves of the tile along p (the controls were tested per tile).
__device__ __forceinline__ void pass_swap(float* lre, float* lim, uint32_t t, const PassBlock& B) {
    const uint32_t tsz = 1u << t, bit = B.pbit, swb = B.rowsw[0];
    const uint32_t lo = bit - 1u;
    for (uint32_t k = threadIdx.x; k < tsz / 2; k += kBlock) {
        const uint32_t i = ((k & ~lo) << 1) | (k & lo);   // bit p of i is 0
        const uint32_t a = lds_sw(i), b = a ^ swb;
        const float r = lre[a], m = lim[a];
        lre[a] = lre[b];
        lim[a] = lim[b];
        lre[b] = r;
        lim[b] = m;
    }
    __syncthreads();
}

// A Pauli rotation exp(i phi P) on the tile as pair updates (B.rot):
// psi'(y) = rc psi(y) + rk (-1)^{|(y ^ rflip) & rpar|} psi(y ^ rflip), with
// rk = i^(ny+1) sin phi; a diagonal string (no flipped bit) scales each
// amplitude by rc + rk (-1)^{|y & rpar|}.  Z qubits of the string outside the
// tile add the same parity to every amplitude of the tile: zodd negates rk.
__device__ __forceinline__ void pass_rot(float* lre, float* lim, uint32_t t, const PassBlock& B, bool zodd) {
    const uint32_t tsz = 1u << t;
    const float c = B.rc, kr = zodd ? -B.rkr : B.rkr, ki = zodd ? -B.rki : B.rki;
    if (B.pbit == 0) {
        for (uint32_t y = threadIdx.x; y < tsz; y += kBlock) {
            const bool odd = (__popc(y & B.rpar) & 1u) != 0u;
            const float fr = odd ? c - kr : c + kr, fi = odd ? -ki : ki;
            const uint32_t a = lds_sw(y);
            const float r = lre[a], m = lim[a];
            lre[a] = r * fr - m * fi;
            lim[a] = r * fi + m * fr;
        }
    } else {
        const uint32_t lo = B.pbit - 1u, swf = lds_sw(B.rflip);
        for (uint32_t k = threadIdx.x; k < tsz / 2; k += kBlock) {
            const uint32_t y0 = ((k & ~lo) << 1) | (k & lo), y1 = y0 ^ B.rflip;   // pivot bit of y0 is 0
            const uint32_t a0 = lds_sw(y0), a1 = a0 ^ swf;
            const float ar = lre[a0], ai = lim[a0], br = lre[a1], bi = lim[a1];
            const float s0 = (__popc(y0 & B.rpar) & 1u) ? -1.f : 1.f;
            const float s1 = (__popc(y1 & B.rpar) & 1u) ? -1.f : 1.f;
            const float tbr = s1 * (kr * br - ki * bi), tbi = s1 * (kr * bi + ki * br);
            const float tar = s0 * (kr * ar - ki * ai), tai = s0 * (kr * ai + ki * ar);
            lre[a0] = c * ar + tbr;
            lim[a0] = c * ai + tbi;
            lre[a1] = c * br + tar;
            lim[a1] = c * bi + tai;
        }
    }
    __syncthreads();
}

// All MFMA blocks of a launch have K = the fusion width; K = 0 marks a swap
// or a sparse rotation.
template <int K>
__global__ __launch_bounds__(kBlock) void k_bpass(float* __restrict__ re, float* __restrict__ im, const PassArgs a) {
    extern __shared__ float lds[];
    const uint32_t tsz = 1u << a.t;
    float* lre = lds;
    float* lim = lds + tsz;
    const uint64_t tile = ((uint64_t)(blockIdx.x & 7u) * (gridDim.x >> 3) + (blockIdx.x >> 3));
    if (tile >= a.ntiles) return;      // whole workgroup leaves together
    const uint64_t base = deposit(tile, a.freemask) | a.fixval;
    const uint32_t tid = threadIdx.x;
    // tile bits 0, 1 are qubits 0, 1: 4 consecutive amplitudes, one aligned group
    // in LDS.  Chunk i of a thread is element 4 tid + 1024 i; every load is
    // issued before the LDS writes.
    constexpr uint32_t kCh = (1u << kPassTileBits) / 4 / kBlock;
    const uint64_t toff = base | deposit(4u * tid, a.tmask);
    const uint64_t ibit0 = deposit(4u * kBlock, a.tmask), ibit1 = deposit(8u * kBlock, a.tmask);
    auto chunk_off = [&](uint32_t i) { return ((i & 1u) ? ibit0 : 0ull) | ((i & 2u) ? ibit1 : 0ull); };
    floatx4 vr[kCh], vi[kCh];
#pragma unroll
    for (uint32_t i = 0; i < kCh; ++i)
        if (tid + i * kBlock < tsz / 4) {
            const uint64_t g = toff | chunk_off(i);
            vr[i] = __builtin_nontemporal_load(reinterpret_cast<const floatx4*>(re + g));
            vi[i] = __builtin_nontemporal_load(reinterpret_cast<const floatx4*>(im + g));
        }
#pragma unroll
    for (uint32_t i = 0; i < kCh; ++i)
        if (tid + i * kBlock < tsz / 4) {
            const uint32_t sw = lds_sw(4 * (tid + i * kBlock));
            *reinterpret_cast<floatx4*>(lre + (sw & ~3u)) = perm4(vr[i], sw & 3u);
            *reinterpret_cast<floatx4*>(lim + (sw & ~3u)) = perm4(vi[i], sw & 3u);
        }
    __syncthreads();
    for (uint32_t bi = 0; bi < a.nblocks; ++bi) {
        const PassBlock& B = a.b[bi];
        if ((base & B.mo) != B.vo) continue;   // uniform: a control of this tile is 0
        if (B.K != 0) pass_block<K>(lre, lim, a.t, B);
        else if (B.rot) pass_rot(lre, lim, a.t, B, (__popcll(base & B.zo) & 1u) != 0u);
        else pass_swap(lre, lim, a.t, B);
    }
    // the product of the scalars that select this tile, in FP64, rounded once
    const bool scale = a.nscal != 0;
    float fr = 1.f, fi = 0.f;
    if (scale) {
        double pr = 1.0, pi = 0.0;
        for (uint32_t k = 0; k < a.nscal; ++k)
            if ((base & a.sc[k].mo) == a.sc[k].vo) {
                const double r = pr * a.sc[k].re - pi * a.sc[k].im;
                pi = pr * a.sc[k].im + pi * a.sc[k].re;
                pr = r;
            }
        fr = (float)pr;
        fi = (float)pi;
    }
#pragma unroll
    for (uint32_t i = 0; i < kCh; ++i)
        if (tid + i * kBlock < tsz / 4) {
            const uint64_t g = toff | chunk_off(i);
            const uint32_t sw = lds_sw(4 * (tid + i * kBlock));
            floatx4 vr = perm4(*reinterpret_cast<const floatx4*>(lre + (sw & ~3u)), sw & 3u);
            floatx4 vi = perm4(*reinterpret_cast<const floatx4*>(lim + (sw & ~3u)), sw & 3u);
            if (scale) {
                const floatx4 nr = vr * fr - vi * fi;
                vi = vr * fi + vi * fr;
                vr = nr;
            }
            __builtin_nontemporal_store(vr, reinterpret_cast<floatx4*>(re + g));
            __builtin_nontemporal_store(vi, reinterpret_cast<floatx4*>(im + g));
        }
}

__global__ __launch_bounds__(kBlock) void k_init32(float* __restrict__ d, uint64_t n4) {
    const uint64_t it = item_id();
    if (it >= n4) return;
    floatx4 v = {0.f, 0.f, 0.f, 0.f};
    if (it == 0) v.x = 1.f;
    reinterpret_cast<floatx4*>(d)[it] = v;
}

// FP64 sums of |psi|^2 over (i & M) == V, split by bit `split` (0: one sum).
// Lanes 0-31 read re, 32-63 im; a wave covers 32 x 4 amplitudes.
__global__ __launch_bounds__(kBlock) void k_norm32(const float* __restrict__ re, const float* __restrict__ im,
                                                   uint64_t n4, uint64_t M, uint64_t V, uint64_t split,
                                                   double* __restrict__ partials) {
    __shared__ double sh[2][kBlock / 64];
    const uint64_t t = item_id();
    const int lane = threadIdx.x & 63;
    const uint64_t it = (t >> 6) * 32 + (uint64_t)(lane & 31);
    double s0 = 0.0, s1 = 0.0;
    if (it < n4) {
        const floatx4 v = reinterpret_cast<const floatx4*>(lane >= 32 ? im : re)[it];
        const uint64_t i = it * 4;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const uint64_t j = i + (uint64_t)e;
            if ((j & M) != V) continue;
            const double x = (double)v[e];
            if (split && (j & split)) s1 += x * x;
            else s0 += x * x;
        }
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = s0;
        sh[1][threadIdx.x >> 6] = s1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a0 = 0.0, a1 = 0.0;
        for (int k = 0; k < kBlock / 64; ++k) {
            a0 += sh[0][k];
            a1 += sh[1][k];
        }
        partials[2 * blockIdx.x] = a0;
        partials[2 * blockIdx.x + 1] = a1;
    }
}

// <ref|psi> in FP64: sum conj(r) p over all amplitudes (ref FP64, psi FP32).
__global__ __launch_bounds__(kBlock) void k_overlap(const double* __restrict__ rre, const double* __restrict__ rim,
                                                    const float* __restrict__ pre, const float* __restrict__ pim,
                                                    uint64_t n2, double* __restrict__ partials) {
    __shared__ double sh[2][kBlock / 64];
    const uint64_t it = item_id();
    double sr = 0.0, si = 0.0;
    if (it < n2) {
        const dv2f ar = reinterpret_cast<const dv2f*>(rre)[it], ai = reinterpret_cast<const dv2f*>(rim)[it];
        const float2 br = reinterpret_cast<const float2*>(pre)[it], bi = reinterpret_cast<const float2*>(pim)[it];
        sr = ar.x * (double)br.x + ai.x * (double)bi.x + ar.y * (double)br.y + ai.y * (double)bi.y;
        si = ar.x * (double)bi.x - ai.x * (double)br.x + ar.y * (double)bi.y - ai.y * (double)br.y;
    }
    sr = wave_sum(sr);
    si = wave_sum(si);
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = sr;
        sh[1][threadIdx.x >> 6] = si;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a0 = 0.0, a1 = 0.0;
        for (int k = 0; k < kBlock / 64; ++k) {
            a0 += sh[0][k];
            a1 += sh[1][k];
        }
        partials[2 * blockIdx.x] = a0;
        partials[2 * blockIdx.x + 1] = a1;
    }
}

__global__ __launch_bounds__(kBlock) void k_sum_pairs32(const double* __restrict__ in, uint64_t npairs, uint64_t chunk,
                                                        double* __restrict__ out) {
    __shared__ double sh[2][kBlock / 64];
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
        sh[0][threadIdx.x >> 6] = s0;
        sh[1][threadIdx.x >> 6] = s1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = 0.0, t1 = 0.0;
        for (int k = 0; k < kBlock / 64; ++k) {
            t0 += sh[0][k];
            t1 += sh[1][k];
        }
        out[2 * blockIdx.x] = t0;
        out[2 * blockIdx.x + 1] = t1;
    }
}

// Collapse on bit q == keep and rescale by the FP64 factor.
__global__ __launch_bounds__(kBlock) void k_collapse32(float* __restrict__ re, float* __restrict__ im, uint64_t n4,
                                                       uint64_t qbit, uint64_t keepbit, double scale) {
    const uint64_t t = item_id();
    const int lane = threadIdx.x & 63;
    const uint64_t it = (t >> 6) * 32 + (uint64_t)(lane & 31);
    if (it >= n4) return;
    floatx4* p = reinterpret_cast<floatx4*>(lane >= 32 ? im : re) + it;
    floatx4 v = *p;
    const uint64_t i = it * 4;
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = (((i + (uint64_t)e) & qbit) == keepbit) ? (float)((double)v[e] * scale) : 0.f;
    *p = v;
}

constexpr uint64_t kPartChunk = 2048;

// Two-level fixed-order reduction of `g` partial pairs into d_result[0..1].
hipError_t finish_sum(suqa_hip_f32* c, uint32_t g) {
    const double* in = c->d_partials;
    uint64_t np = g;
    if (np > kPartChunk) {
        const uint64_t g2 = (np + kPartChunk - 1) / kPartChunk;
        double* out = c->d_partials + 2 * (uint64_t)g;
        hipLaunchKernelGGL(k_sum_pairs32, dim3((uint32_t)g2), dim3(kBlock), 0, c->stream, in, np, kPartChunk, out);
        in = out;
        np = g2;
    }
    hipLaunchKernelGGL(k_sum_pairs32, dim3(1), dim3(kBlock), 0, c->stream, in, np, np, c->d_result);
    return hipGetLastError();
}

int read2(suqa_hip_f32* c) {
    SQF_HIP(hipMemcpyAsync(c->h_result, c->d_result, 2 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SQF_HIP(hipStreamSynchronize(c->stream));
    return SUQA_OK;
}

int norm_split(suqa_hip_f32* c, uint64_t M, uint64_t V, uint64_t split) {
    const uint64_t n4 = c->N / 4;
    const uint32_t g = grid_for(((n4 + 31) / 32) * 64);
    hipLaunchKernelGGL(k_norm32, dim3(g), dim3(kBlock), 0, c->stream, c->d_re, c->d_im, n4, M, V, split,
                       c->d_partials);
    SQF_HIP(hipGetLastError());
    SQF_HIP(finish_sum(c, g));
    return read2(c);
}

// Stacked real form [[Mr, -Mi], [Mi, Mr]] in MFMA A-operand order:
// out[(rb * KS + kk) * 64 + lane] = A[rb*32 + lane%32][2 kk + lane/32].
void mfma_order(const Block& b, float* out) {
    const uint32_t Dc = 1u << b.K, D = 2 * Dc, KS = D / 2, RB = D / 32;
    auto A = [&](uint32_t i, uint32_t k) -> double {
        const bool ii = i >= Dc, ki = k >= Dc;
        const std::complex<double> m = b.m[(size_t)(i % Dc) * Dc + (k % Dc)];
        if (!ii && !ki) return m.real();
        if (!ii && ki) return -m.imag();
        if (ii && !ki) return m.imag();
        return m.real();
    };
    for (uint32_t rb = 0; rb < RB; ++rb)
        for (uint32_t kk = 0; kk < KS; ++kk)
            for (uint32_t l = 0; l < 64; ++l) out[((size_t)rb * KS + kk) * 64 + l] = (float)A(rb * 32 + l % 32, 2 * kk + l / 32);
}

// Host: tile geometry and MFMA-ordered matrices of one block.
void block_args(const Block& b, uint32_t nq, BlockArgs& a) {
    std::memset(&a, 0, sizeof(a));
    uint64_t bm = 0;
    for (uint32_t j = 0; j < b.K; ++j) bm |= 1ull << b.qubits[j];
    std::vector<uint32_t> low;
    for (uint32_t q = 0; q < nq && low.size() < (size_t)(tile_bits1(b.K) - b.K); ++q)
        if (!(((bm | b.cmask) >> q) & 1ull)) low.push_back(q);
    a.L = (uint32_t)low.size();
    uint64_t tm = bm;
    for (uint32_t q : low) tm |= 1ull << q;
    a.tmask = tm;
    a.ntile = (uint32_t)__builtin_popcountll(tm);
    auto cpos = [&](uint32_t q) {   // compact position of qubit q: tile bits below it
        return (uint32_t)__builtin_popcountll(tm & ((1ull << q) - 1ull));
    };
    for (uint32_t r = 0; r < (1u << b.K); ++r) {
        uint32_t o = 0;
        for (uint32_t j = 0; j < b.K; ++j)
            if ((r >> j) & 1u) o |= 1u << cpos(b.qubits[j]);
        a.rowoff[r] = o;
    }
    for (uint32_t cidx = 0; cidx < (1u << a.L); ++cidx) {
        uint32_t o = 0;
        for (uint32_t j = 0; j < a.L; ++j)
            if ((cidx >> j) & 1u) o |= 1u << cpos(low[j]);
        a.coloff[cidx] = o;
    }
    a.freemask = ((1ull << nq) - 1ull) & ~(tm | b.cmask);
    a.fixval = b.cmask;
    a.ntiles = 1ull << __builtin_popcountll(a.freemask);
}

// Uploads the MFMA-ordered matrices of blocks [lo, hi) and launches them.
int apply_chunk(suqa_hip_f32* c, const std::vector<Block>& blocks, size_t lo, size_t hi) {
    std::vector<size_t> off(hi - lo);
    size_t tot = 0;
    for (size_t i = lo; i < hi; ++i) {
        off[i - lo] = tot;
        tot += (size_t)(2u << blocks[i].K) * (2u << blocks[i].K);
    }
    std::vector<float> h(tot);
    for (size_t i = lo; i < hi; ++i) mfma_order(blocks[i], h.data() + off[i - lo]);
    if (c->mats_cap < tot) {
        if (c->d_mats) SQF_HIP(hipFree(c->d_mats));
        c->d_mats = nullptr;
        c->mats_cap = 0;
        void* p = nullptr;
        SQF_HIP(hipMalloc(&p, tot * sizeof(float)));
        c->d_mats = static_cast<float*>(p);
        c->mats_cap = tot;
    }
    SQF_HIP(hipMemcpyAsync(c->d_mats, h.data(), tot * sizeof(float), hipMemcpyHostToDevice, c->stream));
    for (size_t i = lo; i < hi; ++i) {
        BlockArgs a;
        block_args(blocks[i], c->nq, a);
        if (a.L < 5) return fail(SUQA_ERR_ARG, "state too small for this block width and its controls");
        a.A = c->d_mats + off[i - lo];
        const uint32_t g = (uint32_t)(((a.ntiles + 7) / 8) * 8);
        const size_t lds = sizeof(float) * 2 * ((size_t)1 << a.ntile);
        const double amps = (double)a.ntiles * (double)(1ull << a.ntile);
        const int tk = timing_begin(c, "block" + std::to_string(blocks[i].K), 16.0 * amps,
                                    8.0 * (double)(1u << blocks[i].K) * amps);
        switch (blocks[i].K) {
        case 4: hipLaunchKernelGGL(k_block<4>, dim3(g), dim3(kBlock), lds, c->stream, c->d_re, c->d_im, a); break;
        case 5: hipLaunchKernelGGL(k_block<5>, dim3(g), dim3(kBlock), lds, c->stream, c->d_re, c->d_im, a); break;
        case 6: hipLaunchKernelGGL(k_block<6>, dim3(g), dim3(kBlock), lds, c->stream, c->d_re, c->d_im, a); break;
        default: return fail(SUQA_ERR_ARG, "block width must be 4..6");
        }
        timing_end(c, tk);
        SQF_HIP(hipGetLastError());
    }
    // the host matrix buffer must outlive the copy, and the device buffer is
    // rewritten by the next chunk
    SQF_HIP(hipStreamSynchronize(c->stream));
    timing_collect(c);
    return SUQA_OK;
}

// Applies the blocks in order, 256 at a time (at most 16 MiB of matrices).
int apply_blocks(suqa_hip_f32* c, const std::vector<Block>& blocks) {
    constexpr size_t kChunk = 256;
    for (size_t lo = 0; lo < blocks.size(); lo += kChunk)
        if (int rc = apply_chunk(c, blocks, lo, std::min(blocks.size(), lo + kChunk))) return rc;
    return SUQA_OK;
}

// ---- block passes (host) ----------------------------------------------------

struct PassLaunch {
    PassArgs a;
    size_t off[kPassBlocks];   // each block's matrix in the chunk buffer (floats)
};

// Uploads the chunk buffer, points the launches at it, launches them in order
// and waits (the buffer is rewritten by the next chunk).
int flush_passes(suqa_hip_f32* c, std::vector<PassLaunch>& L, std::vector<float>& buf) {
    if (L.empty()) return SUQA_OK;
    if (c->mats_cap < buf.size()) {
        if (c->d_mats) SQF_HIP(hipFree(c->d_mats));
        c->d_mats = nullptr;
        c->mats_cap = 0;
        void* p = nullptr;
        SQF_HIP(hipMalloc(&p, buf.size() * sizeof(float)));
        c->d_mats = static_cast<float*>(p);
        c->mats_cap = buf.size();
    }
    SQF_HIP(hipMemcpyAsync(c->d_mats, buf.data(), buf.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
    for (auto& l : L) {
        for (uint32_t k = 0; k < l.a.nblocks; ++k) l.a.b[k].A = c->d_mats + l.off[k];
        const uint32_t g = (uint32_t)(((l.a.ntiles + 7) / 8) * 8);
        const size_t lds = sizeof(float) * 2 * ((size_t)1 << l.a.t);
        const double amps = (double)l.a.ntiles * (double)(1ull << l.a.t);
        uint32_t dense = 0;   // MFMA blocks of the pass (swaps and sparse rotations use none)
        for (uint32_t k = 0; k < l.a.nblocks; ++k) dense += l.a.b[k].K ? 1u : 0u;
        const int tk = timing_begin(c, "bpass" + std::to_string(l.a.K), 16.0 * amps,
                                    8.0 * (double)(1u << l.a.K) * amps * dense);
        switch (l.a.K) {
        case 4: hipLaunchKernelGGL(k_bpass<4>, dim3(g), dim3(kBlock), lds, c->stream, c->d_re, c->d_im, l.a); break;
        case 5: hipLaunchKernelGGL(k_bpass<5>, dim3(g), dim3(kBlock), lds, c->stream, c->d_re, c->d_im, l.a); break;
        case 6: hipLaunchKernelGGL(k_bpass<6>, dim3(g), dim3(kBlock), lds, c->stream, c->d_re, c->d_im, l.a); break;
        default: return fail(SUQA_ERR_ARG, "block width must be 4..6");
        }
        timing_end(c, tk);
        SQF_HIP(hipGetLastError());
    }
    SQF_HIP(hipStreamSynchronize(c->stream));
    timing_collect(c);
    L.clear();
    buf.clear();
    return SUQA_OK;
}

// One launch of blocks [lo, hi) of a pass on tile qubits T.
int add_launch(suqa_hip_f32* c, const std::vector<Block>& blocks, size_t lo, size_t hi, uint64_t T, uint32_t t,
               uint32_t width, const std::vector<PassScal>* scal, std::vector<PassLaunch>& L,
               std::vector<float>& buf) {
    const uint32_t nq = c->nq;
    PassLaunch l;
    std::memset(&l, 0, sizeof(l));
    PassArgs& a = l.a;
    a.t = t;
    a.tmask = T;
    a.K = width;
    if (scal) {
        a.nscal = (uint32_t)scal->size();
        for (uint32_t k = 0; k < a.nscal; ++k) a.sc[k] = (*scal)[k];
    }
    uint32_t k = 0;
    uint32_t pos[64] = {};   // tile position of each tile qubit
    for (uint32_t q = 0; q < nq; ++q)
        if ((T >> q) & 1ull) pos[q] = k++;
    uint64_t mcom = ~0ull;
    a.nblocks = (uint32_t)(hi - lo);
    for (size_t i = lo; i < hi; ++i) {
        const Block& b = blocks[i];
        PassBlock& B = a.b[i - lo];
        B.mo = B.vo = b.cmask;
        mcom &= b.cmask;
        if (b.swap) {
            B.K = 0;
            B.pbit = 1u << pos[b.qubits[0]];
            B.rowsw[0] = lds_sw(B.pbit);
            l.off[i - lo] = 0;
            continue;
        }
        if (b.rot) {   // qubit masks -> tile positions
            auto tile_bits = [&](uint64_t m) {
                uint32_t r = 0;
                for (uint32_t q = 0; q < nq; ++q)
                    if ((m >> q) & 1ull) r |= 1u << pos[q];
                return r;
            };
            B.K = 0;
            B.rot = 1;
            B.rflip = tile_bits(b.rot->rflip);   // flipped qubits are tile qubits
            B.rpar = tile_bits(b.rot->rpar & T);
            B.zo = b.rot->rpar & ~T;               // Z qubits kept outside the tile
            B.pbit = B.rflip & (0u - B.rflip);
            B.rc = (float)b.rot->rc;
            B.rkr = (float)b.rot->rk.real();
            B.rki = (float)b.rot->rk.imag();
            l.off[i - lo] = 0;
            continue;
        }
        B.K = b.K;
        uint32_t inblock = 0;
        for (uint32_t j = 0; j < b.K; ++j) {
            B.rowsw[j] = lds_sw(1u << pos[b.qubits[j]]);
            inblock |= 1u << pos[b.qubits[j]];
        }
        // column bits 0..4 (the lanes): non-block positions whose bank vectors
        // raise the GF(2) rank, lowest first; then the other positions ascending
        uint32_t nc = 0, taken = 0, basis[5] = {}, rank = 0;
        for (uint32_t p = 0; p < t && nc < 5; ++p) {
            if ((inblock >> p) & 1u) continue;
            uint32_t v = kSwVecAt(p);
            for (uint32_t r = 0; r < rank; ++r) v = std::min(v, v ^ basis[r]);
            if (!v) continue;
            basis[rank++] = v;
            std::sort(basis, basis + rank, [](uint32_t x, uint32_t y) { return x > y; });
            taken |= 1u << p;
            B.colsw[nc++] = lds_sw(1u << p);
        }
        for (uint32_t p = 0; p < t; ++p)
            if (!(((inblock | taken) >> p) & 1u)) B.colsw[nc++] = lds_sw(1u << p);
        // MFMA-ordered stacked matrix, then the tile offsets of the block rows
        const size_t D = (size_t)2 << b.K;
        l.off[i - lo] = buf.size();
        buf.resize(buf.size() + D * D);
        mfma_order(b, buf.data() + l.off[i - lo]);
    }
    // controls shared by every block: tiles where one of them is 0 are not
    // visited (unless the launch applies scalars, which every tile may need)
    if (a.nblocks == 0 || a.nscal != 0) mcom = 0;
    a.freemask = ((1ull << nq) - 1ull) & ~(T | mcom);
    a.fixval = mcom;
    a.ntiles = 1ull << __builtin_popcountll(a.freemask);
    L.push_back(l);
    return SUQA_OK;
}

// Applies ops in order: cut into passes of consecutive ops whose targets fit a
// tile of t qubits, each pass fused into width-qubit blocks inside its tile
// and applied by k_bpass, several blocks per HBM round trip.  The tile is
// settled for the whole pass first, from what each op must flip: every qubit
// of a dense op, the target of a controlled X, the flipped qubits of a sparse
// rotation; a one-qubit diagonal needs none.  Then whatever of an op lies
// outside the tile stays there: the controls of a controlled X become its
// block's control mask, a one-qubit diagonal becomes per-tile scalars, and a
// sparse rotation's Z qubits add a parity per tile.  None of these qubits is
// flipped in the pass (flipped qubits are tile qubits), so the kept-out parts
// commute with everything else in it.  The controls an op comes with
// (Op::cmask, the gc_mask of a Pauli op) always stay outside the tile.
int apply_ops(suqa_hip_f32* c, const std::vector<PassOp>& items, uint32_t width, uint32_t* nblocks) {
    const uint32_t nq = c->nq;
    const uint32_t t = std::min<uint32_t>(kPassTileBits, nq);
    if (t < width + 5) return fail(SUQA_ERR_ARG, "state too small for this block width");
    std::vector<PassLaunch> L;
    std::vector<float> buf;
    uint32_t total = 0;
    size_t i = 0;
    while (i < items.size()) {
        // 1. the pass and its required tile qubits.  Qubits 0..4 (128-byte runs
        //    of FP32 amplitudes) unless the first op comes with controls there.
        uint64_t T = 0;
        for (uint32_t q = 0; q < nq && __builtin_popcountll(T) < 5; ++q)
            if (!((items[i].op.cmask >> q) & 1ull)) T |= 1ull << q;
        uint64_t out = 0;    // controls the ops come with: never tile qubits
        uint64_t refd = 0;   // qubits the ops may leave outside (controls, diagonals, Z)
        uint32_t nscal = 0;  // scalars needed at most (diagonals outside the tile so far)
        size_t j = i;
        while (j < items.size() && j - i < 64) {
            const PassOp& it = items[j];
            const Op& op = it.op;
            uint64_t need = 0, opt = 0;
            uint32_t ns = 0;
            if (it.diag) {
                opt = 1ull << it.dq;
                if (!((T >> it.dq) & 1ull)) ns = (it.d0 != 1.0 ? 1u : 0u) + (it.d1 != 1.0 ? 1u : 0u);
            } else if (it.ctrl) {
                need = 1ull << it.tgt;
                opt = it.ctrl;
            } else if (op.rot) {
                need = op.rflip;
                opt = op.rpar & ~op.rflip;
            } else {
                for (uint32_t k2 = 0; k2 < op.n; ++k2) need |= 1ull << op.qs[k2];
            }
            if ((op.cmask & (T | need)) != 0 || (need & out) != 0 || (uint32_t)__builtin_popcountll(T | need) > t ||
                nscal + ns > kPassScal)
                break;
            T |= need;
            out |= op.cmask;
            refd |= opt;
            nscal += ns;
            ++j;
        }
        if (j == i) return fail(SUQA_ERR_ARG, "an op does not fit a tile");
        // the tile loads move 4 consecutive amplitudes: qubits 0 and 1 must be tile bits
        if ((T & 3ull) != 3ull) return fail(SUQA_ERR_ARG, "controls on qubits 0 or 1 are not supported by FP32 passes");
        // 2. pad the tile: first qubits no op of the pass refers to, then the others
        //    (a control, diagonal or Z qubit inside the tile joins its op's block)
        for (uint32_t q = 0; q < nq && (uint32_t)__builtin_popcountll(T) < t; ++q)
            if (!(((out | refd) >> q) & 1ull)) T |= 1ull << q;
        for (uint32_t q = 0; q < nq && (uint32_t)__builtin_popcountll(T) < t; ++q)
            if (!((out >> q) & 1ull)) T |= 1ull << q;
        if ((uint32_t)__builtin_popcountll(T) < t) return fail(SUQA_ERR_ARG, "state too small for a tile besides the controls");
        // 3. the ops against the settled tile
        std::vector<Op> part;
        std::vector<PassScal> scal;
        part.reserve(j - i);
        for (size_t k2 = i; k2 < j; ++k2) {
            const PassOp& it = items[k2];
            if (it.diag && !((T >> it.dq) & 1ull)) {
                const uint64_t bq = 1ull << it.dq;
                if (it.d0 != 1.0) scal.push_back(PassScal{bq, 0ull, it.d0.real(), it.d0.imag()});
                if (it.d1 != 1.0) scal.push_back(PassScal{bq, bq, it.d1.real(), it.d1.imag()});
                continue;
            }
            Op op = it.op;
            if (it.ctrl && (it.ctrl & ~T)) {
                const uint64_t inside = it.ctrl & T;
                op = cx_op(inside, it.tgt);
                op.cmask = it.ctrl & ~T;
                op.swap = inside == 0;   // X on the target alone: a swap in LDS
            }
            part.push_back(std::move(op));
        }
        const auto blocks = fuse(part, nq, width, 256, T);
        total += (uint32_t)blocks.size();
        // the scalars ride on the pass's last launch (they commute with its blocks)
        for (size_t lo = 0; lo < blocks.size() || (lo == 0 && !scal.empty()); lo += kPassBlocks) {
            const size_t hi = std::min(blocks.size(), lo + kPassBlocks);
            add_launch(c, blocks, lo, hi, T, t, width, hi == blocks.size() ? &scal : nullptr, L, buf);
            if (buf.size() > (size_t)4 << 20)   // 16 MiB of matrices per upload
                if (int rc = flush_passes(c, L, buf)) return rc;
            if (blocks.empty()) break;
        }
        i = j;
    }
    if (nblocks) *nblocks = total;
    return flush_passes(c, L, buf);
}

// Applies ops in the context's mode: block passes, or the whole list fused
// into blocks applied one launch each.
int run_ops(suqa_hip_f32* c, const std::vector<PassOp>& items, uint32_t width, uint32_t* nblocks) {
    if (c->passes) return apply_ops(c, items, width, nblocks);
    std::vector<Op> ops;
    ops.reserve(items.size());
    for (const PassOp& it : items) ops.push_back(it.op);
    const auto blocks = fuse(ops, c->nq, width, 256);
    if (nblocks) *nblocks = (uint32_t)blocks.size();
    return apply_blocks(c, blocks);
}

std::vector<PassOp> as_pass_ops(std::vector<Op>&& ops) {
    std::vector<PassOp> v(ops.size());
    for (size_t k = 0; k < ops.size(); ++k) v[k].op = std::move(ops[k]);
    return v;
}

// Mode 3 (gate passes): the gates go through the FP64 engine's gate entry
// points on a gate queue whose state is this float state, so they become the
// same pair and phase sweeps, and its pass scheduler applies them as float
// tile passes (tiled.hip, k_tile<float>): every gate is computed on its own,
// in FP32 VALU arithmetic, with no fused matrices.  Each call applies what it
// queued before it returns.
int gate_queue(suqa_hip_f32* c, suqa_hip_ctx** out) {
    if (!c->gq) {
        c->gq = sqe::f32_gate_queue(c->device, c->stream, c->nq, c->d_re, c->d_im);
        if (!c->gq) return fail(SUQA_ERR_NOMEM, "gate queue");
    }
    c->gq->tile_bits = std::min(c->gate_tile_bits, c->nq);
    c->gq->timing = c->timing;
    *out = c->gq;
    return SUQA_OK;
}
// rc: what the entry point returned (its message is suqa_hip_last_error's)
int gate_queue_done(suqa_hip_f32* c, int rc, uint32_t* npasses, uint64_t passes_before) {
    if (rc != SUQA_OK) {
        c->gq->queue.clear();
        return fail(rc, suqa_hip_last_error());
    }
    const hipError_t e = sqe::flush(c->gq);
    if (e != hipSuccess) {
        c->gq->queue.clear();
        const std::string why = suqa_hip_last_error();
        return fail(SUQA_ERR_HIP, std::string("gate pass: ") + hipGetErrorString(e) + (why.empty() ? "" : " (" + why + ")"));
    }
    if (npasses) *npasses = (uint32_t)(c->gq->passes - passes_before);
    return SUQA_OK;
}

// Workload gates must name distinct in-range qubits (a block's tile geometry
// is built from them).
int check_gates(const suqa_hip_f32* c, const suqa_hip_gate* gates, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        const suqa_hip_gate& g = gates[i];
        if (g.target >= c->nq || g.kind < 0 || g.kind > SUQA_GATE_MCX || g.nctrl > 3) return fail(SUQA_ERR_ARG, "bad gate");
        uint64_t seen = 1ull << g.target;
        if (g.kind == SUQA_GATE_CX || g.kind == SUQA_GATE_MCX)
            for (uint32_t k = 0; k < g.nctrl; ++k) {
                if (g.ctrl[k] >= c->nq || ((seen >> g.ctrl[k]) & 1ull)) return fail(SUQA_ERR_ARG, "bad control");
                seen |= 1ull << g.ctrl[k];
            }
    }
    return SUQA_OK;
}

bool width_ok(uint32_t nq, uint32_t K) { return K >= (uint32_t)kMinK && K <= (uint32_t)kMaxK && nq >= K + 5; }
// mode 3 takes any width (it fuses nothing)
bool width_ok(const suqa_hip_f32* c, uint32_t K) { return c->passes == 3 || width_ok(c->nq, K); }

}  // namespace
}  // namespace sqf

using namespace sqf;

extern "C" {

const char* suqa_hip_f32_last_error(void) { return sqf::g_err.c_str(); }

int suqa_hip_f32_setup(uint32_t nq, suqa_hip_f32** out) {
    if (!out) return fail(SUQA_ERR_ARG, "null output");
    *out = nullptr;
    if (nq < 9 || nq > 33) return fail(SUQA_ERR_ARG, "the FP32 engine needs 9 <= nq <= 33");
    auto* c = new suqa_hip_f32();
    c->nq = nq;
    c->N = 1ull << nq;
    void* p = nullptr;
    auto cleanup = [&](int rc) {
        suqa_hip_f32_clear(c);
        return rc;
    };
    if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess)
        return cleanup(fail(SUQA_ERR_HIP, "stream creation failed"));
    if (hipMalloc(&p, 2 * c->N * sizeof(float)) != hipSuccess) return cleanup(fail(SUQA_ERR_NOMEM, "state allocation failed"));
    c->d_data = static_cast<float*>(p);
    c->d_re = c->d_data;
    c->d_im = c->d_data + c->N;
    c->partial_cap = c->N / 128 + 4096;
    if (hipMalloc(&p, 2 * c->partial_cap * sizeof(double)) != hipSuccess)
        return cleanup(fail(SUQA_ERR_NOMEM, "partials allocation failed"));
    c->d_partials = static_cast<double*>(p);
    if (hipMalloc(&p, 4 * sizeof(double)) != hipSuccess) return cleanup(fail(SUQA_ERR_NOMEM, "result allocation failed"));
    c->d_result = static_cast<double*>(p);
    if (hipHostMalloc(&p, 4 * sizeof(double), hipHostMallocDefault) != hipSuccess)
        return cleanup(fail(SUQA_ERR_NOMEM, "pinned allocation failed"));
    c->h_result = static_cast<double*>(p);
    int rc = suqa_hip_f32_init_state(c);
    if (rc) return cleanup(rc);
    *out = c;
    return SUQA_OK;
}

int suqa_hip_f32_clear(suqa_hip_f32* c) {
    if (!c) return SUQA_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto& r : c->recs) {
        (void)hipEventDestroy(r.e0);
        (void)hipEventDestroy(r.e1);
    }
    for (auto e : c->evpool) (void)hipEventDestroy(e);
    sqe::f32_gate_queue_free(c->gq);
    if (c->d_data) (void)hipFree(c->d_data);
    if (c->d_partials) (void)hipFree(c->d_partials);
    if (c->d_result) (void)hipFree(c->d_result);
    if (c->h_result) (void)hipHostFree(c->h_result);
    if (c->d_mats) (void)hipFree(c->d_mats);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return SUQA_OK;
}

int suqa_hip_f32_init_state(suqa_hip_f32* c) {
    if (!c || !c->d_data) return fail(SUQA_ERR_ARG, "null context");
    const uint64_t n4 = 2 * c->N / 4;
    hipLaunchKernelGGL(k_init32, dim3(grid_for(n4)), dim3(kBlock), 0, c->stream, c->d_data, n4);
    SQF_HIP(hipGetLastError());
    return SUQA_OK;
}

int suqa_hip_f32_run_gates(suqa_hip_f32* c, const suqa_hip_gate* gates, uint32_t ngates, uint32_t width,
                           uint32_t* nblocks) {
    if (!c || !c->d_data || (ngates && !gates)) return fail(SUQA_ERR_ARG, "null argument");
    if (!width_ok(c, width)) return fail(SUQA_ERR_ARG, "fusion width must be 4..6 and nq >= width + 5");
    if (int rc = check_gates(c, gates, ngates)) return rc;
    if (c->passes == 3) {
        suqa_hip_ctx* q = nullptr;
        if (int rc = gate_queue(c, &q)) return rc;
        const uint64_t p0 = q->passes;
        return gate_queue_done(c, suqa_hip_run_gates(q, gates, ngates), nblocks, p0);
    }
    std::vector<PassOp> items;
    items.reserve(ngates);
    for (uint32_t i = 0; i < ngates; ++i) items.push_back(pass_gate(gates[i]));
    return run_ops(c, items, width, nblocks);
}

int suqa_hip_f32_run_paulis(suqa_hip_f32* c, const suqa_hip_pauli_op* ops, uint32_t nops, uint32_t width,
                            uint32_t* nblocks) {
    if (!c || !c->d_data || (nops && !ops)) return fail(SUQA_ERR_ARG, "null argument");
    if (!width_ok(c, width)) return fail(SUQA_ERR_ARG, "fusion width must be 4..6 and nq >= width + 5");
    if (c->passes == 3) {
        suqa_hip_ctx* q = nullptr;
        if (int rc = gate_queue(c, &q)) return rc;
        const uint64_t p0 = q->passes;
        return gate_queue_done(c, suqa_hip_run_paulis(q, ops, nops), nblocks, p0);
    }
    std::vector<Op> list(nops);
    for (uint32_t i = 0; i < nops; ++i) {
        if (!pauli_op(ops[i], list[i])) return fail(SUQA_ERR_ARG, "malformed Pauli op, a control on one of its qubits, or more than 6 qubits");
        if (list[i].n > width) return fail(SUQA_ERR_ARG, "Pauli op wider than the fusion width");
        if (list[i].n && list[i].qs[list[i].n - 1] >= c->nq) return fail(SUQA_ERR_ARG, "qubit index out of range");
        if (list[i].cmask >> c->nq) return fail(SUQA_ERR_ARG, "control qubit out of range");
        // tiles move 4 consecutive amplitudes: qubits 0 and 1 are always tile qubits
        if (list[i].cmask & 3ull) return fail(SUQA_ERR_ARG, "controls on qubits 0 or 1 are not supported by the FP32 engine");
        if (c->passes == 2 && list[i].n) pauli_rot(ops[i], list[i]);
    }
    return run_ops(c, as_pass_ops(std::move(list)), width, nblocks);
}

int suqa_hip_f32_qft_inverse(suqa_hip_f32* c, const uint32_t* qs, uint32_t n, uint32_t width) {
    if (!c || !c->d_data || (n && !qs)) return fail(SUQA_ERR_ARG, "null argument");
    if (!width_ok(c, width)) return fail(SUQA_ERR_ARG, "fusion width must be 4..6 and nq >= width + 5");
    uint64_t seen = 0;
    for (uint32_t k = 0; k < n; ++k) {
        if (qs[k] >= c->nq || ((seen >> qs[k]) & 1ull)) return fail(SUQA_ERR_ARG, "bad register");
        seen |= 1ull << qs[k];
    }
    if (c->passes == 3) {
        suqa_hip_ctx* q = nullptr;
        if (int rc = gate_queue(c, &q)) return rc;
        return gate_queue_done(c, suqa_hip_apply_qft_inverse(q, qs, n, 0), nullptr, q->passes);
    }
    std::vector<Op> list;
    qft_inverse_ops(qs, n, list);
    return run_ops(c, as_pass_ops(std::move(list)), width, nullptr);
}

int suqa_hip_f32_set_passes(suqa_hip_f32* c, uint32_t on) {
    if (!c) return fail(SUQA_ERR_ARG, "null context");
    if (on > 3) return fail(SUQA_ERR_ARG, "mode must be 0 (one block per launch), 1 (block passes), 2 (block passes, sparse Pauli rotations) or 3 (gate passes)");
    c->passes = on;
    return SUQA_OK;
}

int suqa_hip_f32_get_passes(const suqa_hip_f32* c, uint32_t* on) {
    if (!c || !on) return fail(SUQA_ERR_ARG, "null argument");
    *on = c->passes;
    return SUQA_OK;
}

int suqa_hip_f32_autotune(suqa_hip_f32* c, const suqa_hip_gate* gates, uint32_t ngates, uint32_t* best_width,
                          double* ms_per_gate) {
    if (!c || !c->d_data || !best_width || !gates || !ngates) return fail(SUQA_ERR_ARG, "null argument or empty sample");
    if (int rc = check_gates(c, gates, ngates)) return rc;
    std::vector<PassOp> ops;
    for (uint32_t i = 0; i < ngates; ++i) ops.push_back(pass_gate(gates[i]));
    const std::vector<PassOp> warm(ops.begin(), ops.begin() + std::min<size_t>(ops.size(), 8));
    const uint32_t nwarm = std::min<uint32_t>(ngates, 8);
    double best = 1e300, best_block = 1e300;
    uint32_t best_mode = 0, best_tb = c->gate_tile_bits;
    *best_width = 0;
    // modes 0 and 1 at widths 4..6 (slots 0..5), mode 3 at tile bits 12, 13 (slots 6, 7)
    for (uint32_t slot = 0; slot < 8; ++slot) {
        const uint32_t mode = slot < 3 ? 0u : slot < 6 ? 1u : 3u;
        const uint32_t K = mode == 3 ? (uint32_t)kMinK : kMinK + slot % 3;
        const uint32_t tb = mode == 3 ? 12u + (slot - 6u) : c->gate_tile_bits;
        if (ms_per_gate) ms_per_gate[slot] = -1.0;
        if (!width_ok(c->nq, K) || (mode == 3 && tb > c->nq)) continue;
        c->passes = mode;
        const uint32_t keep_tb = c->gate_tile_bits;
        c->gate_tile_bits = tb;
        // warm-up on the first ops, then time the whole sample
        int rc = mode == 3 ? suqa_hip_f32_run_gates(c, gates, nwarm, K, nullptr) : run_ops(c, warm, K, nullptr);
        const auto t0 = std::chrono::steady_clock::now();
        if (!rc) rc = mode == 3 ? suqa_hip_f32_run_gates(c, gates, ngates, K, nullptr) : run_ops(c, ops, K, nullptr);
        c->gate_tile_bits = keep_tb;
        if (rc) return rc;
        SQF_HIP(hipStreamSynchronize(c->stream));
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        const double per = ms / std::max<uint32_t>(ngates, 1);
        if (ms_per_gate) ms_per_gate[slot] = per;
        if (mode != 3 && per < best_block) {
            best_block = per;
            *best_width = K;
        }
        if (per < best) {
            best = per;
            best_mode = mode;
            best_tb = tb;
        }
    }
    c->passes = best_mode;
    c->gate_tile_bits = best_tb;
    return *best_width ? SUQA_OK : fail(SUQA_ERR_ARG, "no fusion width fits this state");
}

int suqa_hip_f32_set_tile_bits(suqa_hip_f32* c, uint32_t bits) {
    if (!c) return fail(SUQA_ERR_ARG, "null context");
    if (bits < 12 || bits > sqe::kMaxTileBits32) return fail(SUQA_ERR_ARG, "gate-pass tile bits must be 12 or 13");
    c->gate_tile_bits = bits;
    return SUQA_OK;
}

int suqa_hip_f32_get_tile_bits(const suqa_hip_f32* c, uint32_t* bits) {
    if (!c || !bits) return fail(SUQA_ERR_ARG, "null argument");
    *bits = c->gate_tile_bits;
    return SUQA_OK;
}

int suqa_hip_f32_vnorm(suqa_hip_f32* c, double* norm) {
    if (!c || !c->d_data) return fail(SUQA_ERR_ARG, "null context");
    int rc = norm_split(c, 0, 0, 0);
    if (rc) return rc;
    if (norm) *norm = std::sqrt(c->h_result[0]);
    return SUQA_OK;
}

int suqa_hip_f32_prob_filter(suqa_hip_f32* c, const uint32_t* qs, const uint32_t* q_mask, uint32_t n, double* prob) {
    if (!c || !c->d_data || (n && (!qs || !q_mask))) return fail(SUQA_ERR_ARG, "null argument");
    uint64_t M = 0, V = 0;
    for (uint32_t k = 0; k < n; ++k) {
        if (qs[k] >= c->nq) return fail(SUQA_ERR_ARG, "qubit index out of range");
        M |= 1ull << qs[k];
        if (q_mask[k]) V |= (uint64_t)q_mask[k] << qs[k];
    }
    if (V & ~M) {
        if (prob) *prob = 0.0;
        return SUQA_OK;
    }
    int rc = norm_split(c, M, V, 0);
    if (rc) return rc;
    if (prob) *prob = c->h_result[0];
    return SUQA_OK;
}

int suqa_hip_f32_measure_qbit(suqa_hip_f32* c, uint32_t q, double rdoub, uint32_t* cout) {
    if (!c || !c->d_data) return fail(SUQA_ERR_ARG, "null context");
    if (q >= c->nq) return fail(SUQA_ERR_ARG, "qubit index out of range");
    const uint64_t qb = 1ull << q;
    int rc = norm_split(c, 0, 0, qb);
    if (rc) return rc;
    const double p0 = c->h_result[0], p1 = c->h_result[1];
    const uint32_t cc = (uint32_t)(rdoub < p1);
    const double scale = 1. / std::sqrt(cc ? p1 : p0);
    const uint64_t n4 = c->N / 4;
    hipLaunchKernelGGL(k_collapse32, dim3(grid_for(((n4 + 31) / 32) * 64)), dim3(kBlock), 0, c->stream, c->d_re, c->d_im,
                       n4, qb, cc ? qb : 0ull, scale);
    SQF_HIP(hipGetLastError());
    if (cout) *cout = cc;
    return SUQA_OK;
}

int suqa_hip_f32_fidelity(suqa_hip_f32* c, suqa_hip_ctx* ref, double* fidelity) {
    if (!c || !c->d_data || !ref) return fail(SUQA_ERR_ARG, "null argument");
    uint32_t rnq = 0;
    uint64_t rN = 0;
    double *rre = nullptr, *rim = nullptr;
    if (suqa_hip_state_info(ref, &rnq, &rN, &rre, &rim) != SUQA_OK || rnq != c->nq)
        return fail(SUQA_ERR_ARG, "reference state has another size");
    SQF_HIP(hipDeviceSynchronize());   // the reference context runs on its own stream
    const uint64_t n2 = c->N / 2;
    const uint32_t g = grid_for(n2);
    hipLaunchKernelGGL(k_overlap, dim3(g), dim3(kBlock), 0, c->stream, rre, rim, c->d_re, c->d_im, n2, c->d_partials);
    SQF_HIP(hipGetLastError());
    SQF_HIP(finish_sum(c, g));
    int rc = read2(c);
    if (rc) return rc;
    const double ov = c->h_result[0] * c->h_result[0] + c->h_result[1] * c->h_result[1];
    // normalised by both norms (FP64 sums): rounding drift of the FP32 norm does
    // not count as fidelity
    double n32 = 0.0, n64 = 0.0;
    rc = suqa_hip_f32_vnorm(c, &n32);
    if (rc) return rc;
    if (suqa_hip_vnorm(ref, &n64) != SUQA_OK) return fail(SUQA_ERR_HIP, "reference norm failed");
    if (fidelity) *fidelity = ov / (n32 * n32 * n64 * n64);
    return SUQA_OK;
}

int suqa_hip_f32_download(suqa_hip_f32* c, float* re, float* im) {
    if (!c || !c->d_data) return fail(SUQA_ERR_ARG, "null context");
    if (re) SQF_HIP(hipMemcpyAsync(re, c->d_re, c->N * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (im) SQF_HIP(hipMemcpyAsync(im, c->d_im, c->N * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    SQF_HIP(hipStreamSynchronize(c->stream));
    return SUQA_OK;
}

int suqa_hip_f32_timing_enable(suqa_hip_f32* c, int on) {
    if (!c) return fail(SUQA_ERR_ARG, "null context");
    c->timing = on != 0;
    if (c->gq) c->gq->timing = c->timing;
    return SUQA_OK;
}

int suqa_hip_f32_timing_reset(suqa_hip_f32* c) {
    if (!c) return fail(SUQA_ERR_ARG, "null context");
    SQF_HIP(hipStreamSynchronize(c->stream));
    timing_collect(c);
    c->stats.clear();
    if (c->gq) {
        sqe::timing_flush(c->gq);
        c->gq->stats.clear();
    }
    return SUQA_OK;
}

int suqa_hip_f32_timing_report(suqa_hip_f32* c, suqa_hip_kstat* out, uint32_t max, uint32_t* n) {
    if (!c || (max && !out)) return fail(SUQA_ERR_ARG, "null argument");
    SQF_HIP(hipStreamSynchronize(c->stream));
    timing_collect(c);
    std::vector<suqa_hip_f32::Stat> all = c->stats;
    if (c->gq) {   // the gate passes of mode 3 (VALU: no MFMA flops)
        sqe::timing_flush(c->gq);
        for (const auto& g : c->gq->stats) {
            all.emplace_back();
            all.back().name = g.name;
            all.back().launches = g.launches;
            all.back().ms = g.total_ms;
            all.back().bytes = g.alg_bytes;
        }
    }
    uint32_t k = 0;
    for (; k < all.size() && k < max; ++k) {
        std::memset(&out[k], 0, sizeof(out[k]));
        std::strncpy(out[k].name, all[k].name.c_str(), sizeof(out[k].name) - 1);
        out[k].launches = all[k].launches;
        out[k].total_ms = all[k].ms;
        out[k].alg_bytes = all[k].bytes;
        out[k].flops = all[k].flops;
    }
    if (n) *n = (uint32_t)all.size();
    return SUQA_OK;
}

int suqa_hip_f32_synchronize(suqa_hip_f32* c) {
    if (!c || !c->stream) return fail(SUQA_ERR_ARG, "null context");
    SQF_HIP(hipStreamSynchronize(c->stream));
    return SUQA_OK;
}

}  // extern "C"
