// noalloc.hip -- suqa::prob_filter on a state the caller holds, with no
// device allocation: the reduction partials live in a fixed device array.
//
// For callers that cannot allocate device memory (code under HIP graph
// capture, serving loops, harnesses that forbid hipMalloc).  The kernels are
// those of suqa_hip_prob_filter (k_norm + the fixed k_sum_pairs tree), so the
// result is bit-identical to it for the same state.
//
// The array `scratch` is one per GPU, held until the process ends.  Host code
// never uses its name as an address: its device address is looked up once per
// device (hipGetSymbolAddress, with that device current) and that pointer is
// what the kernels receive.  Every call writes the partials it reads, so
// nothing needs clearing; calls that overlap on one GPU (other streams, other
// threads) share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "engine.hpp"

namespace {

// States of at most 2^kNoAllocMaxQubits amplitudes: 2 * partial_pairs_capacity
// doubles of partials (4 MiB at 26 qubits) and the 2-double result.
constexpr uint32_t kNoAllocMaxQubits = 26;
constexpr uint64_t kPartialDoubles = 2ull * ((1ull << kNoAllocMaxQubits) / sqe::kBlock +
                                             ((1ull << kNoAllocMaxQubits) / sqe::kBlock + sqe::kPartChunk - 1) /
                                                 sqe::kPartChunk +
                                             8);
constexpr uint64_t kScratchDoubles = kPartialDoubles + 2;

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};

// the message goes where suqa_hip_last_error() reads it
int fail(int code, const std::string& msg) {
    sqe::set_last_error(msg);
    return code;
}

__global__ void k_store1(double* dst, double v) { *dst = v; }
__global__ void k_copy1(double* dst, const double* src) { *dst = *src; }

// The device address of `scratch` on the current device, looked up once.
int scratch_ptr(double** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return fail(SUQA_ERR_HIP, "no current HIP device");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(SUQA_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_ptr[dev];
    return SUQA_OK;
}

}  // namespace

extern "C" {

int suqa_hip_prob_filter_noalloc(const double* d_re, const double* d_im, uint32_t nq, const uint32_t* qs,
                                 const uint32_t* q_mask, uint32_t n, double* d_prob, void* stream) {
    if (!d_re || !d_im || !d_prob || (n && (!qs || !q_mask))) return fail(SUQA_ERR_ARG, "null argument");
    if (nq < 1 || nq > kNoAllocMaxQubits) return fail(SUQA_ERR_ARG, "nq must be in [1, 26] (scratch size)");
    if (sqe::partial_pairs_capacity(1ull << nq) * 2 > kPartialDoubles)
        return fail(SUQA_ERR_ARG, "reduction partials exceed the scratch array");
    uint32_t mqs = 0u, mask = 0u;   // suqa.cu:706-713
    if (!sqe::prob_filter_masks(qs, q_mask, n, &mqs, &mask)) return fail(SUQA_ERR_ARG, "qubit index out of range");
    double* ptr = nullptr;
    if (int rc = scratch_ptr(&ptr)) return rc;
    const hipStream_t st = static_cast<hipStream_t>(stream);
    sqe::Sel sel;
    if (!sqe::make_sel(nq, mqs, mask, &sel)) {   // empty selection: probability 0
        hipLaunchKernelGGL(k_store1, dim3(1), dim3(1), 0, st, d_prob, 0.0);
        const hipError_t e = hipGetLastError();
        return e == hipSuccess ? SUQA_OK : fail(SUQA_ERR_HIP, hipGetErrorString(e));
    }
    double* result = ptr + kPartialDoubles;
    hipError_t e = sqe::launch_norm_tree(d_re, d_im, nq, sel, 0, ptr, result, st, nullptr, "prob_filter", 0.0);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_copy1, dim3(1), dim3(1), 0, st, d_prob, result);
        e = hipGetLastError();
    }
    return e == hipSuccess ? SUQA_OK : fail(SUQA_ERR_HIP, hipGetErrorString(e));
}

}  // extern "C"
