// qms_batch.hip -- thousands of independent QMS chains in a few launches.
//
// A chain's state is 2^(r+4) amplitudes: 3 system qubits, the r-qubit
// register slot at qubit 3 and the acceptance qubit on top.  With E_old in
// its own register (two registers, 2r + 4 qubits in the reference layout) the
// state is stored in the same 2^(r+4) amplitudes, exactly: the E_old register
// reads one definite value on every nonzero amplitude except during its own
// phase estimation, and there E_new and the acceptance qubit read 0.  So the
// slot holds E_old while it is phase-estimated and measured, and E_new after;
// the E_old value is kept in a scalar (vold), every gate that reads it (W)
// takes the scalar, and a measurement of one of its qubits (a reset) is a
// norm and a scale of the stored amplitudes (measure_fixed).  Every other
// amplitude of the full state is exactly zero and stays zero under every
// gate the chain applies, so the stored ones are computed exactly as the
// full-state chain computes them.
//
// Two layouts of those amplitudes, with one control flow (Chain<Ops>):
// * RegOps, batches of FP64 chains no larger than the CU count
//   (latency first): a workgroup of N threads per chain, one amplitude per
//   thread.  A gate exchanges each amplitude with its partner lane (a shuffle
//   inside the wavefront, LDS and a barrier across wavefronts); a measurement
//   is a block reduction.
// * LdsOps, everything else (throughput): 1 to 8 wavefronts per chain, the
//   fewer chains LDS holds per CU the more, taking chains from a work queue in
//   device memory, the state in the workgroup's LDS, every gate a short loop
//   over the workgroup's lanes.
//
// The control flow is metro_step (qms.cuh:373-545) with the same pcg32 draw
// order; gate arithmetic keeps the reference's operation order
// (suqa_cpu.hpp:49-66, 150-206; qms.cuh:286-307) and device code is built
// with -ffp-contract=off, so a chain reproduces the reference CPU path's
// accept/reject sequence for its seed.  Measurement sums are wave
// reductions (another summation order than the reference's sequential
// loop: amplitudes agree to ~1e-15, and outcomes are identical unless a
// uniform draw falls within that distance of a probability).
//
// FP32 mode (the precision pattern): the same kernels over float amplitudes,
// with every measurement probability summed and the collapse scale computed
// in FP64.  Its chains follow the FP64 ones until a
// uniform draw falls within the FP32 rounding of a probability.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "qms_batch.hpp"
#include "suqa_hip.h"
#include "suqa_qms.h"

namespace qb {
namespace {

constexpr double kTwoSqInv = 0.7071067811865475;   // suqa.cuh:57
constexpr int E0 = 3;                              // first energy qubit (qms.cuh:79-88)

// The value of lane id ^ m for m = 1, 2, 4, 8 (DPP moves inside a row of 16
// lanes: xor 1, 2 are quad permutes; xor 4 = mirror within 8, then xor 3;
// xor 8 = mirror within 16, then within 8), on both dwords of a double.
template <int C1, int C2>
__device__ __forceinline__ double dpp_xor(double v) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    int lo = (int)(uint32_t)u, hi = (int)(uint32_t)(u >> 32);
    lo = __builtin_amdgcn_mov_dpp(lo, C1, 0xF, 0xF, true);
    hi = __builtin_amdgcn_mov_dpp(hi, C1, 0xF, 0xF, true);
    if constexpr (C2 >= 0) {
        lo = __builtin_amdgcn_mov_dpp(lo, C2, 0xF, 0xF, true);
        hi = __builtin_amdgcn_mov_dpp(hi, C2, 0xF, 0xF, true);
    }
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// Butterfly sum over the 64 lanes (every lane ends with a sum of all; lane
// 0's is the one used): DPP moves for xor 1..8, the permlane swaps for 16 and
// 32 (the swap of rows, or of halves, hands each lane its partner's value
// beside its own; 12 two-dword LDS permutes before).
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_xor<0xB1, -1>(v);
    v += dpp_xor<0x4E, -1>(v);
    v += dpp_xor<0x141, 0x1B>(v);
    v += dpp_xor<0x140, 0x141>(v);
    {
        const uint64_t u = (uint64_t)__double_as_longlong(v);
        const auto l = __builtin_amdgcn_permlane16_swap((uint32_t)u, (uint32_t)u, false, false);
        const auto h = __builtin_amdgcn_permlane16_swap((uint32_t)(u >> 32), (uint32_t)(u >> 32), false, false);
        v = __longlong_as_double((long long)(((uint64_t)h[0] << 32) | l[0])) +
            __longlong_as_double((long long)(((uint64_t)h[1] << 32) | l[1]));
    }
    {
        const uint64_t u = (uint64_t)__double_as_longlong(v);
        const auto l = __builtin_amdgcn_permlane32_swap((uint32_t)u, (uint32_t)u, false, false);
        const auto h = __builtin_amdgcn_permlane32_swap((uint32_t)(u >> 32), (uint32_t)(u >> 32), false, false);
        v = __longlong_as_double((long long)(((uint64_t)h[0] << 32) | l[0])) +
            __longlong_as_double((long long)(((uint64_t)h[1] << 32) | l[1]));
    }
    return v;
}

__device__ __forceinline__ uint32_t ins0(uint32_t p, uint32_t q) {   // insert a zero bit at q
    return ((p >> q) << (q + 1)) | (p & ((1u << q) - 1u));
}

// pcg32 of the `pcg` wrapper (pcg32.h:53-68, 121-130; Rand.hpp:257-273)
struct Rng {
    uint64_t state, inc;
    __device__ void seed(uint64_t s) {
        state = 0;
        inc = (1ull << 1u) | 1u;
        next();
        state += s;
        next();
    }
    __device__ uint32_t next() {
        const uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((~rot + 1u) & 31u));
    }
    __device__ double doub() {
        return __longlong_as_double((long long)(((uint64_t)next() << 20) | 0x3ff0000000000000ULL)) - 1.0;
    }
};

// What every chain layout holds.
struct ChainBase {
    uint32_t r, n, N, L, acc;
    uint32_t en;         // first qubit of the register slot (E0)
    bool two;            // E_old in its own register (qms.cuh:72-85, 134-136, 201-202)
    uint32_t vold;       // two registers: the value the E_old register holds (definite, see Chain)
    uint32_t lane;       // 0 on the one thread that writes a chain's outputs
    uint64_t acc_amps;   // LdsOps: amplitude accesses (reads + writes) the gates and measurements must make
    Rng rng;
    const ChainDev* P;
};

// The chain's state in the workgroup's LDS: TW wavefronts per chain (1 to 8,
// more the fewer chains LDS holds per CU), gates as loops over 64 * TW-lane
// strides.
template <typename T, int TW>
struct LdsOps : ChainBase {
    struct alignas(2 * sizeof(T)) Amp {
        T r, i;
    };
    Amp* st;             // LDS: amplitude i at st[i], real and imaginary parts adjacent (one access)
    double* red;         // LDS: the waves' partial sums of a measurement (TW > 1)

    __device__ __forceinline__ void sync() const { __syncthreads(); }

    // ------------------------------------------------------------ gates
    __device__ __forceinline__ void init_zero() {
        acc_amps += N;
        for (uint32_t i = lane; i < N; i += 64 * TW) {
            st[i].r = i == 0 ? T(1) : T(0);
            st[i].i = T(0);
        }
        sync();
    }

    // H on q (suqa_cpu.hpp:186-204) for states with (i & gc) == gc
    __device__ __forceinline__ void h(uint32_t q, uint32_t gc) {
        acc_amps += 4ull * ((N / 2) >> __builtin_popcount(gc));
        for (uint32_t p = lane; p < N / 2; p += 64 * TW) {
            const uint32_t a = ins0(p, q);
            if ((a & gc) != gc) continue;
            const uint32_t b = a | (1u << q);
            const T a0r = st[a].r, a0i = st[a].i, a1r = st[b].r, a1i = st[b].i;
            st[a].r = T(kTwoSqInv) * (a0r + a1r);
            st[a].i = T(kTwoSqInv) * (a0i + a1i);
            st[b].r = T(kTwoSqInv) * (a0r - a1r);
            st[b].i = T(kTwoSqInv) * (a0i - a1i);
        }
        sync();
    }

    // X on qt for states whose control bits read cv (suqa_cpu.hpp:239-251)
    __device__ __forceinline__ void mcx(uint32_t qt, uint32_t cm, uint32_t cv) {
        acc_amps += 4ull * ((N / 2) >> (__builtin_popcount(cm) - 1));
        for (uint32_t p = lane; p < N / 2; p += 64 * TW) {
            const uint32_t a = ins0(p, qt);
            const uint32_t b = a | (1u << qt);
            if ((b & cm) != cv) continue;
            T t = st[a].r;
            st[a].r = st[b].r;
            st[b].r = t;
            t = st[a].i;
            st[a].i = st[b].i;
            st[b].i = t;
        }
        sync();
    }
    __device__ __forceinline__ void x(uint32_t q) { mcx(q, 1u << q, 1u << q); }
    __device__ __forceinline__ void cx(uint32_t qc, uint32_t qt) {
        const uint32_t m = (1u << qc) | (1u << qt);
        mcx(qt, m, m);
    }

    // util_rotate2 on states with all bits of m set (u1 / cu1 with q_mask 1)
    __device__ __forceinline__ void phase(uint32_t q1, uint32_t q2, double cd, double sd) {
        const T c = T(cd), s = T(sd);
        const uint32_t lo = q1 < q2 ? q1 : q2, hi = q1 < q2 ? q2 : q1;
        const uint32_t m = (1u << q1) | (1u << q2);
        const uint32_t cnt = q1 == q2 ? N / 2 : N / 4;
        acc_amps += 2ull * cnt;
        for (uint32_t p = lane; p < cnt; p += 64 * TW) {
            const uint32_t i = (q1 == q2 ? ins0(p, lo) : ins0(ins0(p, lo), hi)) | m;
            const T cpy = st[i].r;
            st[i].r = cpy * c - st[i].i * s;
            st[i].i = st[i].i * c + cpy * s;
        }
        sync();
    }

    // exp(i phi X_a X_b) on states with (i & gc) == gc (suqa_cpu.hpp:422-424)
    __device__ __forceinline__ void xx(uint32_t qa, uint32_t qb_, double cd, double sd, uint32_t gc) {
        const T c = T(cd), s = T(sd);
        const uint32_t pv = qa < qb_ ? qa : qb_;
        const uint32_t X = (1u << qa) | (1u << qb_);
        acc_amps += 4ull * ((N / 2) >> __builtin_popcount(gc));
        for (uint32_t p = lane; p < N / 2; p += 64 * TW) {
            const uint32_t j = ins0(p, pv);
            if ((j & gc) != gc) continue;
            const uint32_t k = j ^ X;
            const T jr = st[j].r, ji = st[j].i, kr = st[k].r, ki = st[k].i;
            st[j].r = jr * c - ki * s;
            st[j].i = ji * c + kr * s;
            st[k].r = kr * c - ji * s;
            st[k].i = ki * c + jr * s;
        }
        sync();
    }

    // measure_qbit (suqa.cu:673-696): c = (rd < P(q=1)), collapse, renormalize
    __device__ __forceinline__ uint32_t measure(uint32_t q, double rd) {
        acc_amps += N + N / 2 + N;   // read all, read the kept half, write all
        double s0 = 0.0, s1 = 0.0;
        for (uint32_t i = lane; i < N; i += 64 * TW) {
            const double a = st[i].r, b = st[i].i;   // FP64 probabilities in both modes
            const double p = a * a + b * b;
            if ((i >> q) & 1u) s1 += p;
            else s0 += p;
        }
        s0 = wave_sum(s0);
        s1 = wave_sum(s1);
        if constexpr (TW > 1) {   // the last use of red ended at a barrier (the collapse below)
            if ((lane & 63u) == 0u) {
                red[2 * (lane >> 6)] = s0;
                red[2 * (lane >> 6) + 1] = s1;
            }
            __syncthreads();
            s0 = 0.0;
            s1 = 0.0;
            for (int w = 0; w < TW; ++w) {
                s0 += red[2 * w];
                s1 += red[2 * w + 1];
            }
        }
        const uint32_t c = __builtin_amdgcn_readfirstlane((uint32_t)(rd < s1));
        const double scale = 1. / sqrt(c ? s1 : s0);
        for (uint32_t i = lane; i < N; i += 64 * TW) {
            const bool keep = ((i >> q) & 1u) == c;
            st[i].r = keep ? T((double)st[i].r * scale) : T(0);
            st[i].i = keep ? T((double)st[i].i * scale) : T(0);
        }
        sync();
        return c;
    }

    __device__ __forceinline__ void reset(uint32_t q, double rd) {   // suqa.cu:730-738
        if (measure(q, rd)) x(q);
    }

    // measure_qbit of a qubit that is not stored and reads bitv on every
    // stored amplitude: P(1) is the whole norm or 0, the stored amplitudes are
    // the kept half and are scaled as measure() scales it (an outcome of
    // probability 0 leaves 0 * 1/sqrt(0) = NaN, as it would in the full state)
    __device__ __forceinline__ uint32_t measure_fixed(uint32_t bitv, double rd) {
        acc_amps += 3ull * N;   // read all twice, write all
        double s = 0.0;
        for (uint32_t i = lane; i < N; i += 64 * TW) {
            const double a = st[i].r, b = st[i].i;
            s += a * a + b * b;
        }
        s = wave_sum(s);
        if constexpr (TW > 1) {   // the last use of red ended at a barrier
            if ((lane & 63u) == 0u) red[2 * (lane >> 6)] = s;
            __syncthreads();
            s = 0.0;
            for (int w = 0; w < TW; ++w) s += red[2 * w];
        }
        const uint32_t c = __builtin_amdgcn_readfirstlane((uint32_t)(rd < (bitv ? s : 0.0)));
        const double f = c == bitv ? 1. / sqrt(s) : __builtin_nan("");
        for (uint32_t i = lane; i < N; i += 64 * TW) {
            st[i].r = T((double)st[i].r * f);
            st[i].i = T((double)st[i].i * f);
        }
        sync();
        return c;
    }

    // the stored amplitudes as FP64 (re, im) pairs (after the last step)
    __device__ __forceinline__ void dump(double* dst) const {
        for (uint32_t i = lane; i < N; i += 64 * TW) {
            dst[2 * i] = (double)st[i].r;
            dst[2 * i + 1] = (double)st[i].i;
        }
    }

    // qms::apply_W, dense form (qms.cuh:286-307); acc is the top qubit, E_old
    // the classical value (two registers: the value the E_old register holds).
    __device__ __forceinline__ void W(uint32_t Eold) {
        const uint32_t half = N / 2;
        acc_amps += 2ull * N;
        for (uint32_t j = lane; j < half; j += 64 * TW) {
            const uint32_t i = j + half;
            const uint32_t En = (i >> en) & (L - 1);
            const double* row = P->wtab + 2 * (size_t)Eold * L;
            const T fs1 = T(row[2 * En]), fs2 = T(row[2 * En + 1]);
            T t = st[j].r;
            const T ri = st[i].r;
            st[j].r = fs2 * t + fs1 * ri;
            st[i].r = fs1 * t - fs2 * ri;
            t = st[j].i;
            const T ii = st[i].i;
            st[j].i = fs2 * t + fs1 * ii;
            st[i].i = fs1 * t - fs2 * ii;
        }
        sync();
    }

};

// One amplitude per thread (N <= 1024 threads, one workgroup per chain):
// a gate exchanges with its partner lane (a shuffle inside the wavefront, LDS
// across wavefronts) and every thread updates its own amplitude with the
// arithmetic of LdsOps, so a gate costs one exchange instead of a loop.  For
// chains that cannot fill the GPU (a single chain): latency, not throughput.
template <typename T>
struct RegOps : ChainBase {
    T vr, vi;            // amplitude idx
    uint32_t idx;
    T* xr;               // LDS exchange buffers (2 x N each, alternating) and
    T* xi;               // reduction partials (2 x 32, alternating): one barrier
    double* red;         // per exchange or reduction, as the previous use of a
    uint32_t xb, rb;     // buffer ended before the barrier of the use in between

    __device__ __forceinline__ void sync() const { __syncthreads(); }

    // amplitude idx ^ m; every thread of the workgroup calls it (m is uniform)
    __device__ __forceinline__ void partner(uint32_t m, T& pr, T& pi) {
        if (m < 64u) {
            pr = __shfl_xor(vr, (int)m, 64);
            pi = __shfl_xor(vi, (int)m, 64);
        } else {
            T* br = xr + xb * N;
            T* bi = xi + xb * N;
            br[idx] = vr;
            bi[idx] = vi;
            __syncthreads();
            pr = br[idx ^ m];
            pi = bi[idx ^ m];
            xb ^= 1u;
        }
    }

    __device__ __forceinline__ void init_zero() {
        vr = idx == 0 ? T(1) : T(0);
        vi = T(0);
    }

    __device__ __forceinline__ void h(uint32_t q, uint32_t gc) {   // suqa_cpu.hpp:186-204
        const uint32_t qb = 1u << q;
        T pr, pi;
        partner(qb, pr, pi);
        if (((idx & ~qb) & gc) != gc) return;
        if (idx & qb) {   // b: kTwoSqInv * (a0 - a1)
            vr = T(kTwoSqInv) * (pr - vr);
            vi = T(kTwoSqInv) * (pi - vi);
        } else {          // a: kTwoSqInv * (a0 + a1)
            vr = T(kTwoSqInv) * (vr + pr);
            vi = T(kTwoSqInv) * (vi + pi);
        }
    }

    __device__ __forceinline__ void mcx(uint32_t qt, uint32_t cm, uint32_t cv) {   // suqa_cpu.hpp:239-251
        const uint32_t qb = 1u << qt;
        T pr, pi;
        partner(qb, pr, pi);
        if (((idx | qb) & cm) != cv) return;
        vr = pr;
        vi = pi;
    }
    __device__ __forceinline__ void x(uint32_t q) { mcx(q, 1u << q, 1u << q); }
    __device__ __forceinline__ void cx(uint32_t qc, uint32_t qt) {
        const uint32_t m = (1u << qc) | (1u << qt);
        mcx(qt, m, m);
    }

    __device__ __forceinline__ void phase(uint32_t q1, uint32_t q2, double cd, double sd) {   // util_rotate2
        const T c = T(cd), s = T(sd);
        const uint32_t m = (1u << q1) | (1u << q2);
        if ((idx & m) != m) return;
        const T cpy = vr;
        vr = cpy * c - vi * s;
        vi = vi * c + cpy * s;
    }

    __device__ __forceinline__ void xx(uint32_t qa, uint32_t qb_, double cd, double sd, uint32_t gc) {   // suqa_cpu.hpp:422-424
        const T c = T(cd), s = T(sd);
        const uint32_t pv = qa < qb_ ? qa : qb_;
        const uint32_t X = (1u << qa) | (1u << qb_);
        T pr, pi;
        partner(X, pr, pi);
        const uint32_t j = (idx >> pv) & 1u ? idx ^ X : idx;
        if ((j & gc) != gc) return;
        const T nr = vr * c - pi * s;   // both ends: own * c -/+ partner's other component * s
        vi = vi * c + pr * s;
        vr = nr;
    }

    __device__ __forceinline__ uint32_t measure(uint32_t q, double rd) {   // suqa.cu:673-696
        const double x = vr, y = vi;   // FP64 probabilities in both precisions
        const double p = x * x + y * y;
        const bool one = (idx >> q) & 1u;
        double s0 = wave_sum(one ? 0.0 : p), s1 = wave_sum(one ? p : 0.0);
        const uint32_t w = idx >> 6, nw = N >> 6;
        double* pr = red + 32 * rb;
        if ((idx & 63u) == 0u) {
            pr[2 * w] = s0;
            pr[2 * w + 1] = s1;
        }
        __syncthreads();
        double t0 = 0.0, t1 = 0.0;
        for (uint32_t k = 0; k < nw; ++k) {
            t0 += pr[2 * k];
            t1 += pr[2 * k + 1];
        }
        rb ^= 1u;
        // every thread read the same partials: uniform, and said so, so that the
        // control flow that depends on it stays scalar
        const uint32_t cbit = __builtin_amdgcn_readfirstlane(rd < t1 ? 1u : 0u);
        const double scale = 1. / sqrt(cbit ? t1 : t0);
        const bool keep = (uint32_t)one == cbit;
        vr = keep ? T((double)vr * scale) : T(0);
        vi = keep ? T((double)vi * scale) : T(0);
        return cbit;
    }

    __device__ __forceinline__ void reset(uint32_t q, double rd) {   // suqa.cu:730-738
        if (measure(q, rd)) x(q);
    }

    // measure_qbit of a qubit that is not stored and reads bitv everywhere
    // (LdsOps::measure_fixed)
    __device__ __forceinline__ uint32_t measure_fixed(uint32_t bitv, double rd) {
        const double x = vr, y = vi;
        const double s = wave_sum(x * x + y * y);
        const uint32_t w = idx >> 6, nw = N >> 6;
        double* pr = red + 32 * rb;
        if ((idx & 63u) == 0u) pr[2 * w] = s;
        __syncthreads();
        double t = 0.0;
        for (uint32_t k = 0; k < nw; ++k) t += pr[2 * k];
        rb ^= 1u;
        const uint32_t c = __builtin_amdgcn_readfirstlane(rd < (bitv ? t : 0.0) ? 1u : 0u);
        const double f = c == bitv ? 1. / sqrt(t) : __builtin_nan("");
        vr = T((double)vr * f);
        vi = T((double)vi * f);
        return c;
    }

    __device__ __forceinline__ void dump(double* dst) const {
        dst[2 * idx] = (double)vr;
        dst[2 * idx + 1] = (double)vi;
    }

    __device__ __forceinline__ void W(uint32_t Eold) {   // qms.cuh:286-307: pairs (j, j + N/2)
        const uint32_t half = N / 2;
        T pr, pi;
        partner(half, pr, pi);
        const uint32_t i = idx | half;
        const uint32_t En = (i >> en) & (L - 1);
        const double* row = P->wtab + 2 * (size_t)Eold * L;
        const T fs1 = T(row[2 * En]), fs2 = T(row[2 * En + 1]);
        if (idx < half) {   // j: fs2 t + fs1 ri
            vr = fs2 * vr + fs1 * pr;
            vi = fs2 * vi + fs1 * pi;
        } else {            // i: fs1 t - fs2 ri
            vr = fs1 * pr - fs2 * vr;
            vi = fs1 * pi - fs2 * vi;
        }
    }
};

// The QMS control flow over either state layout.
template <class Ops>
struct Chain : Ops {
    using Ops::r;
    using Ops::n;
    using Ops::N;
    using Ops::L;
    using Ops::acc;
    using Ops::en;
    using Ops::two;
    using Ops::vold;
    using Ops::lane;
    using Ops::rng;
    using Ops::P;
    using Ops::init_zero;
    using Ops::h;
    using Ops::x;
    using Ops::cx;
    using Ops::phase;
    using Ops::xx;
    using Ops::measure;
    using Ops::reset;
    using Ops::measure_fixed;
    using Ops::W;

    // ------------------------------------------------------- QMS pieces
    __device__ __forceinline__ void system_init() {   // system.cu:6-11
        init_zero();
        vold = 0u;
        x(1);
        h(1, 0u);
        cx(1, 0);
    }

    __device__ __forceinline__ void cevolution(uint32_t base, uint32_t trg, int dir) {   // qms.cuh:106-115, system.cu:21-26
        const uint32_t gc = 1u << (base + trg);
        const double c = P->xx_c[dir][trg], s = P->xx_s[dir][trg];
        xx(1, 2, c, s, gc);
        xx(2, 0, c, s, gc);
        xx(0, 1, c, s, gc);
    }

    __device__ __forceinline__ void Phi() {   // qms.cuh:153-176 on the register slot
        const uint32_t b = en;
        for (uint32_t k = 0; k < r; ++k) h(b + k, 0u);
        for (int trg = (int)r - 1; trg > -1; --trg) {
            cevolution(b, (uint32_t)trg, 0);
            phase(b + trg, b + trg, P->u1_c[0][trg], P->u1_s[0][trg]);
        }
        for (uint32_t o = 0; o < r; ++o) {   // QFT^-1 (suqa.cu:649-661)
            for (uint32_t in = 0; in < o; ++in) phase(b + in, b + o, P->qft_c[1][1 + o - in], P->qft_s[1][1 + o - in]);
            h(b + o, 0u);
        }
    }

    __device__ __forceinline__ void Phi_inverse() {   // qms.cuh:179-198
        for (int o = (int)r - 1; o >= 0; o--) {   // QFT (suqa.cu:635-647)
            h(en + o, 0u);
            for (int in = o - 1; in >= 0; in--) phase(en + in, en + o, P->qft_c[0][1 + o - in], P->qft_s[0][1 + o - in]);
        }
        for (uint32_t trg = 0; trg < r; ++trg) {
            cevolution(en, trg, 1);
            phase(en + trg, en + trg, P->u1_c[1][trg], P->u1_s[1][trg]);
        }
        for (uint32_t k = 0; k < r; ++k) h(en + k, 0u);
    }

    __device__ __forceinline__ void U(uint32_t ci, uint32_t Eold) {   // qms.cuh:335-348
        h(ci, 0u);
        Phi();
        W(Eold);
    }
    __device__ __forceinline__ void U_inverse(uint32_t ci, uint32_t Eold) {   // qms.cuh:350-360
        W(Eold);
        Phi_inverse();
        h(ci, 0u);
    }

    // measure_qbits(bm_enes, cs, extract_rands(r)): the r draws first, then the
    // measurements in order; the draws are replayed from a copy of the
    // generator, so the loop stays rolled and holds no array
    __device__ __forceinline__ uint32_t measure_enes() {
        Rng d = rng;
        for (uint32_t k = 0; k < r; ++k) (void)rng.doub();
        uint32_t v = 0;
        for (uint32_t k = 0; k < r; ++k) v |= measure(en + k, d.doub()) << k;
        return v;
    }

    // r draws, then the resets of the register slot in order
    __device__ __forceinline__ void reset_register() {
        Rng d = rng;
        for (uint32_t k = 0; k < r; ++k) (void)rng.doub();
        for (uint32_t k = 0; k < r; ++k) reset(en + k, d.doub());
    }

    // apply_reset of E_old qubit k (two registers): the qubit reads bit k of
    // vold on every nonzero amplitude, so it is measured by measure_fixed and
    // reads 0 after the reset's X
    __device__ __forceinline__ void reset_old_qubit(uint32_t k, double rd) {
        (void)measure_fixed((vold >> k) & 1u, rd);
        vold &= ~(1u << k);
    }
    __device__ __forceinline__ void reset_old_register() {   // qms.cuh:134-136: r draws, then the resets
        Rng d = rng;
        for (uint32_t k = 0; k < r; ++k) (void)rng.doub();
        for (uint32_t k = 0; k < r; ++k) reset_old_qubit(k, d.doub());
    }

    __device__ __forceinline__ void reset_enes_each_draw() {
        for (uint32_t k = 0; k < r; ++k) reset(en + k, rng.doub());
    }

    __device__ __forceinline__ void measure_rotation() {   // system.cu:76-83
        h(0, 0u);
        h(1, 0u);
        h(2, 0u);
        cx(0, 1);
        cx(0, 2);
        phase(2, 2, P->hp_c, P->hp_s);
    }

    __device__ __forceinline__ double measure_X() {   // system.cu:99-119, op_vals {2,0,2,0,-2,0,-2,0}
        Rng d = rng;   // three draws, then the three measurements
        for (int k = 0; k < 3; ++k) (void)rng.doub();
        uint32_t meas = 0;
        for (int rep = 0; rep < 2; ++rep) {
            measure_rotation();
            if (rep == 0)
                for (uint32_t q = 0; q < 3; ++q) meas |= measure(q, d.doub()) << q;
        }
        return (meas & 1u) ? 0.0 : ((meas & 4u) ? -2.0 : 2.0);
    }

    // metro_step (qms.cuh:373-545); returns the code, E/X through *Ed/*Xv when measured
    __device__ __forceinline__ int metro_step(bool take, uint32_t max_rev, double* Ed, double* Xv) {
        int ret = 0;
        // reset_non_syst_qbits (qms.cuh:128-150): the E_old register first
        if (two) reset_old_register();
        reset_register();
        reset(acc, rng.doub());
        // Phi and the E_old measurement on the register slot.  One register:
        // the measured value is flipped back to 0.  Two registers: E_new and
        // the acceptance qubit read 0 here, so the slot holds E_old during its
        // phase estimation; once measured, E_old is vold and the slot, flipped
        // back to 0, is E_new (the same amplitudes, moved by the same X gates)
        Phi();
        const uint32_t Eold = measure_enes();
        for (uint32_t k = 0; k < r; ++k)
            if ((Eold >> k) & 1u) x(en + k);
        if (two) vold = Eold;
        uint32_t ci;
        {   // draw_C (qms.cuh:207-216, system.cu:123)
            const double e = rng.doub();
            ci = e < 1. / 3 ? 0u : (e < 2. / 3 ? 1u : 2u);
        }
        U(ci, Eold);
        if (measure(acc, rng.doub()) == 1u) {
            const uint32_t Enew = measure_enes();
            Phi_inverse();
            if (!take) return 1;
            *Ed = P->etab[Enew];
            for (uint32_t k = 0; k < r; ++k) {   // two registers: E_old qubit k reset first (qms.cuh:431-434)
                if (two) reset_old_qubit(k, rng.doub());
                if ((Enew >> k) & 1u) x(en + k);
            }
            *Xv = measure_X();
            reset_enes_each_draw();
            Phi();
            (void)measure_enes();
            Phi_inverse();
            return 2;
        }
        U_inverse(ci, Eold);
        uint32_t iters = 0;
        while (iters < max_rev) {
            Phi();
            if (two)   // the E_old readout's r draws (qms.cuh:478-479); the register holds vold
                for (uint32_t k = 0; k < r; ++k) (void)rng.doub();
            const uint32_t Enew = measure_enes();
            Phi_inverse();
            if (Eold == Enew) {
                if (take) {
                    *Ed = P->etab[Eold];
                    reset_enes_each_draw();
                    *Xv = measure_X();
                    reset_enes_each_draw();
                    Phi();
                    (void)measure_enes();
                    Phi_inverse();
                    ret = 4;
                } else {
                    ret = 3;
                }
                break;
            }
            U(ci, Eold);
            (void)measure(acc, rng.doub());
            U_inverse(ci, Eold);
            iters++;
        }
        if (iters == max_rev) return -1;
        return ret;
    }

    // driver loop of qms.cu:132-149
    __device__ __forceinline__ void run(const BatchDev& b, uint32_t chain) {
        rng.seed(P->seed);
        system_init();
        uint32_t s0 = 0, acc_count = 0, nm = 0;
        int32_t* rets = b.rets + (size_t)chain * b.steps;
        uint64_t* rnga = b.rng_after ? b.rng_after + (size_t)chain * b.steps : nullptr;
        double* Eo = b.E + (size_t)chain * b.cap_meas;
        double* Xo = b.X + (size_t)chain * b.cap_meas;
        for (uint32_t s = 0; s < b.steps; ++s) {
            const bool take = (s > s0 + (uint32_t)b.thermalization) && ((s - s0) % b.reset_each == 0u);
            double Ed = 0.0, Xv = 0.0;
            const int ret = metro_step(take, b.max_reverse, &Ed, &Xv);
            if (ret == 2 || ret == 4) {
                if (lane == 0 && nm < b.cap_meas) {
                    Eo[nm] = Ed;
                    Xo[nm] = Xv;
                }
                nm++;
            }
            if (ret < 0) {
                system_init();
                s0 = s + 1;
            }
            if (ret == 1 || ret == 2) acc_count++;
            if (lane == 0) {
                rets[s] = ret;
                if (rnga) rnga[s] = rng.state;
            }
        }
        if (lane == 0) {
            b.n_meas[chain] = nm;
            b.accepted[chain] = acc_count;
            if (b.lds_amps) b.lds_amps[chain] = this->acc_amps;
        }
        if (b.final_state) this->dump(b.final_state + P->state_off);
    }
};

template <typename T, int TW>
__global__ __launch_bounds__(64 * TW) void k_qms_batch(const BatchDev b) {
    extern __shared__ double lds[];
    const uint32_t lane = threadIdx.x;
    for (;;) {
        uint32_t pos = 0;
        if constexpr (TW == 1) {
            if (lane == 0) pos = b.first + atomicAdd(b.next_chain, 1u);
            pos = __builtin_amdgcn_readfirstlane(__shfl(pos, 0, 64));
        } else {   // through LDS; the previous chain's last reads ended at a barrier
            uint32_t* q = reinterpret_cast<uint32_t*>(lds);
            if (lane == 0) *q = b.first + atomicAdd(b.next_chain, 1u);
            __syncthreads();
            pos = __builtin_amdgcn_readfirstlane(*q);
            __syncthreads();
        }
        if (pos >= b.nchains) return;   // every workgroup reaches this once the queue is drained
        const uint32_t chain = b.order[pos];
        Chain<LdsOps<T, TW>> c;
        c.P = &b.chains[chain];
        c.r = c.P->r;
        c.two = b.two_registers != 0;
        c.n = c.r + 4;
        c.N = 1u << c.n;
        c.L = 1u << c.r;
        c.acc = c.n - 1;
        c.en = E0;
        c.lane = lane;
        c.acc_amps = 0;
        c.st = reinterpret_cast<typename LdsOps<T, TW>::Amp*>(lds);
        c.red = reinterpret_cast<double*>(c.st + c.N);
        c.run(b, chain);
        __syncthreads();
    }
}

// One chain per workgroup of N threads, one amplitude per thread (RegOps).
template <typename T>
__global__ __launch_bounds__(1024) void k_qms_chain_reg(const BatchDev b) {
    extern __shared__ double lds[];
    const uint32_t pos = b.first + blockIdx.x;
    if (pos >= b.nchains) return;   // uniform
    const uint32_t chain = b.order[pos];
    Chain<RegOps<T>> c;
    c.P = &b.chains[chain];
    c.r = c.P->r;
    c.two = b.two_registers != 0;
    c.n = c.r + 4;
    c.N = 1u << c.n;
    c.L = 1u << c.r;
    c.acc = c.n - 1;
    c.en = E0;
    c.idx = threadIdx.x;
    c.lane = threadIdx.x;
    c.acc_amps = 0;
    c.red = lds;   // 2 x 32 doubles, then the exchange buffers
    c.xr = reinterpret_cast<T*>(lds + 64);
    c.xi = c.xr + 2 * c.N;
    c.xb = 0;
    c.rb = 0;
    c.run(b, chain);
}

__global__ void k_zero_u32(uint32_t* p) { *p = 0u; }

// ----------------------------------------------------------------- host
void fill_chain(const suqa_qms_chain& in, const suqa_qms_batch_params& p, ChainDev& d, std::vector<double>& wtab) {
    // qms.cu:262-276 and qms.cuh:96-103
    const uint32_t r = in.ene_qbits;
    const uint32_t L = 1u << r;
    const double t_PE_shift = p.ene_min;
    const double t_PE_factor = (L - 1) / (double)(L * (p.ene_max - p.ene_min));
    const double t = t_PE_factor * 8. * atan(1.0);
    const double c_factor = in.beta / (t_PE_factor * L);
    std::memset(&d, 0, sizeof(d));
    d.seed = in.seed;
    d.r = r;
    for (uint32_t trg = 0; trg < r; ++trg) {
        const double powr = (double)(1U << (r - 1 - trg));
        // Phi: cevolution(-powr*t) -> pauli(XX, -(-powr*t)); u1(-powr*t*Emin)
        const double tp = -powr * t;
        sincos(-tp, &d.xx_s[0][trg], &d.xx_c[0][trg]);
        sincos(-powr * t * t_PE_shift, &d.u1_s[0][trg], &d.u1_c[0][trg]);
        // Phi^dag: cevolution(powr*t) -> pauli(XX, -(powr*t)); u1(powr*t*Emin)
        const double ti = powr * t;
        sincos(-ti, &d.xx_s[1][trg], &d.xx_c[1][trg]);
        sincos(powr * t * t_PE_shift, &d.u1_s[1][trg], &d.u1_c[1][trg]);
    }
    // crm phases of the QFT: rphase_m[m] = 2 pi / 2^m (suqa.cu:611-633)
    uint32_t cc = 1;
    for (uint32_t m = 0; m < (uint32_t)kMaxEne + 2; ++m) {
        const double ph = (2.0 * M_PI / (double)cc);
        double neg = ph;
        neg *= -1;
        sincos(ph, &d.qft_s[0][m], &d.qft_c[0][m]);
        sincos(neg, &d.qft_s[1][m], &d.qft_c[1][m]);
        cc <<= 1;
    }
    sincos(M_PI * 0.5, &d.hp_s, &d.hp_c);
    for (uint32_t k = 0; k < L; ++k) d.etab[k] = t_PE_shift + k / (double)(t_PE_factor * L);
    wtab.resize(2 * (size_t)L * L);
    for (uint32_t eo = 0; eo < L; ++eo)
        for (uint32_t en = 0; en < L; ++en) {
            const double dE = (en * c_factor - eo * c_factor);
            double fs1 = 1.0, fs2 = 0.0;
            if (dE > 0) {
                fs1 = exp(-dE / 2.0);
                fs2 = sqrt(1.0 - fs1 * fs1);
            }
            wtab[2 * ((size_t)eo * L + en)] = fs1;
            wtab[2 * ((size_t)eo * L + en) + 1] = fs2;
        }
}

std::string g_err;

}  // namespace
}  // namespace qb

#define QB_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) {                                                             \
            fprintf(stderr, "suqa_qms_batch_run: %s: %s\n", #call, hipGetErrorString(e_));  \
            rc = SUQA_ERR_HIP;                                                              \
            goto done;                                                                      \
        }                                                                                   \
    } while (0)

extern "C" int suqa_qms_batch_run(const suqa_qms_chain* chains, uint32_t nchains, const suqa_qms_batch_params* p,
                                  suqa_qms_batch_result* out) {
    using namespace qb;
    if (!chains || !p || !out || nchains == 0 || !out->rets || !out->n_meas || !out->accepted) return SUQA_ERR_ARG;
    if (p->metro_steps == 0 || p->reset_each == 0 || p->max_reverse == 0 || p->pe_steps <= 0) return SUQA_ERR_ARG;
    if (out->cap_meas && (!out->E || !out->X)) return SUQA_ERR_ARG;
    for (uint32_t c = 0; c < nchains; ++c)
        if (chains[c].ene_qbits < 2 || chains[c].ene_qbits > (uint32_t)kMaxEne || chains[c].seed == 0 ||
            !(chains[c].beta > 0.0))
            return SUQA_ERR_ARG;
    int rc = SUQA_OK;
    std::vector<ChainDev> hc(nchains);
    std::vector<std::vector<double>> wt(nchains);
    size_t wtotal = 0, stotal = 0;
    for (uint32_t c = 0; c < nchains; ++c) {
        fill_chain(chains[c], *p, hc[c], wt[c]);
        wtotal += wt[c].size();
        hc[c].state_off = stotal;
        stotal += 2 * ((size_t)1 << (chains[c].ene_qbits + 4));
    }
    ChainDev* d_chains = nullptr;
    double *d_w = nullptr, *d_E = nullptr, *d_X = nullptr;
    int32_t* d_rets = nullptr;
    uint64_t* d_rng = nullptr;
    uint32_t *d_nm = nullptr, *d_acc = nullptr, *d_next = nullptr, *d_order = nullptr;
    uint64_t* d_lds = nullptr;
    double* d_final = nullptr;
    std::vector<size_t> need(nchains);
    std::vector<uint32_t> order(nchains);
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<hipStream_t> gst;   // a stream per group of equal LDS need: the groups run side by side
    std::vector<hipEvent_t> gev;    // each group's end, waited for on st
    uint32_t ngroups = 0;
    const size_t nsteps = (size_t)nchains * p->metro_steps;
    const size_t nmeas = (size_t)nchains * out->cap_meas;
    BatchDev b{};
    int ncu = 256, lds_max = 0, lds_cu = 0;
    bool reg = false;   // the per-amplitude layout for FP64 one-register chains
    uint32_t grid = 0;
    size_t lds_bytes = 0;
    float ms = 0.f;
    const size_t elt = p->fp32 ? sizeof(float) : sizeof(double);
    QB_HIP(hipStreamCreate(&st));
    QB_HIP(hipMalloc(&d_w, sizeof(double) * wtotal));
    {
        std::vector<double> flat;
        flat.reserve(wtotal);
        for (uint32_t c = 0; c < nchains; ++c) {
            hc[c].wtab = d_w + flat.size();
            flat.insert(flat.end(), wt[c].begin(), wt[c].end());
        }
        QB_HIP(hipMemcpy(d_w, flat.data(), sizeof(double) * flat.size(), hipMemcpyHostToDevice));
    }
    QB_HIP(hipMalloc(&d_chains, sizeof(ChainDev) * nchains));
    QB_HIP(hipMemcpy(d_chains, hc.data(), sizeof(ChainDev) * nchains, hipMemcpyHostToDevice));
    QB_HIP(hipMalloc(&d_rets, sizeof(int32_t) * nsteps));
    if (out->rng_after) QB_HIP(hipMalloc(&d_rng, sizeof(uint64_t) * nsteps));
    QB_HIP(hipMalloc(&d_E, sizeof(double) * (nmeas ? nmeas : 1)));
    QB_HIP(hipMalloc(&d_X, sizeof(double) * (nmeas ? nmeas : 1)));
    QB_HIP(hipMalloc(&d_nm, sizeof(uint32_t) * nchains));
    QB_HIP(hipMalloc(&d_acc, sizeof(uint32_t) * nchains));
    QB_HIP(hipMalloc(&d_next, sizeof(uint32_t) * nchains));   // a work-queue counter per group
    if (out->lds_accesses) QB_HIP(hipMalloc(&d_lds, sizeof(uint64_t) * nchains));
    if (out->final_state) QB_HIP(hipMalloc(&d_final, sizeof(double) * stotal));
    // chains grouped by the LDS their state needs, widest first: when LDS
    // limits the resident waves, one launch per group (the launches run side by
    // side), so narrow chains are not held to the widest chain's occupancy
    for (uint32_t c = 0; c < nchains; ++c) {
        const uint32_t n = chains[c].ene_qbits + 4;   // both layouts (two registers: see Chain)
        need[c] = (p->fp32 ? sizeof(float) : sizeof(double)) * 2 * ((size_t)1 << n);
        order[c] = c;
    }
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return need[x] > need[y]; });
    QB_HIP(hipMalloc(&d_order, sizeof(uint32_t) * nchains));
    QB_HIP(hipMemcpy(d_order, order.data(), sizeof(uint32_t) * nchains, hipMemcpyHostToDevice));
    b.chains = d_chains;
    b.order = d_order;
    b.nchains = nchains;
    b.steps = p->metro_steps;
    b.reset_each = p->reset_each;
    b.max_reverse = p->max_reverse;
    b.thermalization = p->thermalization;
    b.cap_meas = out->cap_meas;
    b.rets = d_rets;
    b.rng_after = d_rng;
    b.E = d_E;
    b.X = d_X;
    b.n_meas = d_nm;
    b.accepted = d_acc;
    b.two_registers = p->two_registers ? 1u : 0u;
    b.lds_amps = d_lds;
    b.final_state = d_final;
    {
        int dev = 0;
        QB_HIP(hipGetDevice(&dev));
        QB_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        QB_HIP(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        QB_HIP(hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    }
    lds_bytes = need[order[0]];   // the widest chain
    {
        if (lds_bytes > (size_t)lds_max) {
            fprintf(stderr, "suqa_qms_batch_run: a chain's state (%zu bytes) exceeds the %d bytes of LDS of a workgroup\n",
                    lds_bytes, lds_max);
            rc = SUQA_ERR_ARG;
            goto done;
        }
    }
    // Batches of no more chains than CUs (latency: a single chain runs 27 k
    // steps/s against 14 k in LDS) of FP64 one-register chains of at most 1024
    // amplitudes run a workgroup per chain with one amplitude per thread
    // (RegOps); larger batches, wider, two-register and FP32 chains run in LDS
    // with 1 to 8 wavefronts per chain (LdsOps: 8.3 against 6.3 M chain-steps/s
    // for 4096 FP64 chains with r in 3..6).
    reg = !p->fp32 && nchains <= (uint32_t)ncu;
    if (!reg || need[order[0]] > 2 * elt * 1024) {
        // One launch sized for the widest chain when every wave of the grid fits
        // in LDS at that size; otherwise one launch per group of equal LDS need.
        const uint64_t waves_cu = (std::min<uint64_t>(nchains, (uint64_t)ncu * 16u) + ncu - 1) / ncu;
        if (waves_cu * lds_bytes <= (uint64_t)lds_cu)
            for (uint32_t c = 0; c < nchains; ++c) need[c] = lds_bytes;
    }
    QB_HIP(hipEventCreate(&e0));
    QB_HIP(hipEventCreate(&e1));
    QB_HIP(hipEventRecord(e0, st));
    // The groups run concurrently, each on its own stream with its own queue
    // (their chains fill the CUs together: every group alone left most wave
    // slots empty while its longest chains finished); st waits for all.
    for (uint32_t lo = 0; lo < nchains;) {   // one wave per chain, LDS sized for the group
        uint32_t hi = lo;
        while (hi < nchains && need[order[hi]] == need[order[lo]]) ++hi;
        b.first = lo;
        b.nchains = hi;
        b.next_chain = d_next + ngroups;
        grid = hi - lo < (uint32_t)ncu * 16u ? hi - lo : (uint32_t)ncu * 16u;
        hipStream_t gs = nullptr;
        hipEvent_t ge = nullptr;
        QB_HIP(hipStreamCreate(&gs));
        gst.push_back(gs);
        QB_HIP(hipEventCreateWithFlags(&ge, hipEventDisableTiming));
        gev.push_back(ge);
        ++ngroups;
        QB_HIP(hipStreamWaitEvent(gs, e0, 0));
        hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, gs, b.next_chain);
        if (reg && need[order[lo]] <= 2 * elt * 1024) {
            const uint32_t Ng = (uint32_t)(need[order[lo]] / (2 * elt));   // amplitudes per chain
            const size_t sh = 64 * sizeof(double) + 4 * (size_t)Ng * elt;
            hipLaunchKernelGGL(k_qms_chain_reg<double>, dim3(hi - lo), dim3(Ng), sh, gs, b);
        } else {
            // waves per chain: the fewer chains LDS holds per CU, the more waves
            // each (up to 8), so the SIMDs stay busy; the measurement's partial
            // sums follow the state
            const uint64_t per_cu = lds_cu / need[order[lo]];
            const int tw = per_cu >= 32 ? 1 : (per_cu >= 16 ? 2 : (per_cu >= 4 ? 4 : 8));
            const size_t sh = need[order[lo]] + (tw > 1 ? 2 * sizeof(double) * tw : 0);
            if (sh > (size_t)lds_max) {
                fprintf(stderr, "suqa_qms_batch_run: %zu bytes of LDS per chain exceed the %d of a workgroup\n", sh,
                        lds_max);
                rc = SUQA_ERR_ARG;
                goto done;
            }
#define QB_LAUNCH_LDS(TYPE, TWV) \
    hipLaunchKernelGGL((k_qms_batch<TYPE, TWV>), dim3(grid), dim3(64 * TWV), sh, gs, b)
            if (p->fp32) {
                if (tw == 1) QB_LAUNCH_LDS(float, 1);
                else if (tw == 2) QB_LAUNCH_LDS(float, 2);
                else if (tw == 4) QB_LAUNCH_LDS(float, 4);
                else QB_LAUNCH_LDS(float, 8);
            } else {
                if (tw == 1) QB_LAUNCH_LDS(double, 1);
                else if (tw == 2) QB_LAUNCH_LDS(double, 2);
                else if (tw == 4) QB_LAUNCH_LDS(double, 4);
                else QB_LAUNCH_LDS(double, 8);
            }
#undef QB_LAUNCH_LDS
        }
        QB_HIP(hipGetLastError());
        QB_HIP(hipEventRecord(ge, gs));
        QB_HIP(hipStreamWaitEvent(st, ge, 0));
        lo = hi;
    }
    QB_HIP(hipEventRecord(e1, st));
    QB_HIP(hipStreamSynchronize(st));
    QB_HIP(hipEventElapsedTime(&ms, e0, e1));
    out->seconds = ms * 1e-3;
    QB_HIP(hipMemcpy(out->rets, d_rets, sizeof(int32_t) * nsteps, hipMemcpyDeviceToHost));
    if (out->rng_after) QB_HIP(hipMemcpy(out->rng_after, d_rng, sizeof(uint64_t) * nsteps, hipMemcpyDeviceToHost));
    if (nmeas) {
        QB_HIP(hipMemcpy(out->E, d_E, sizeof(double) * nmeas, hipMemcpyDeviceToHost));
        QB_HIP(hipMemcpy(out->X, d_X, sizeof(double) * nmeas, hipMemcpyDeviceToHost));
    }
    QB_HIP(hipMemcpy(out->n_meas, d_nm, sizeof(uint32_t) * nchains, hipMemcpyDeviceToHost));
    QB_HIP(hipMemcpy(out->accepted, d_acc, sizeof(uint32_t) * nchains, hipMemcpyDeviceToHost));
    if (d_lds) {   // chains of the per-amplitude layout count nothing (0)
        QB_HIP(hipMemcpy(out->lds_accesses, d_lds, sizeof(uint64_t) * nchains, hipMemcpyDeviceToHost));
        for (uint32_t c = 0; c < nchains; ++c) out->lds_accesses[c] *= 2 * elt;   // accesses -> bytes
    }
    if (d_final) QB_HIP(hipMemcpy(out->final_state, d_final, sizeof(double) * stotal, hipMemcpyDeviceToHost));
done:
    for (hipStream_t gs : gst) (void)hipStreamSynchronize(gs);
    for (hipStream_t gs : gst) (void)hipStreamDestroy(gs);
    for (hipEvent_t ge : gev) (void)hipEventDestroy(ge);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamSynchronize(st);
    (void)hipFree(d_chains);
    (void)hipFree(d_w);
    (void)hipFree(d_E);
    (void)hipFree(d_X);
    (void)hipFree(d_rets);
    (void)hipFree(d_rng);
    (void)hipFree(d_nm);
    (void)hipFree(d_acc);
    (void)hipFree(d_next);
    (void)hipFree(d_order);
    (void)hipFree(d_lds);
    (void)hipFree(d_final);
    if (st) (void)hipStreamDestroy(st);
    return rc;
}
