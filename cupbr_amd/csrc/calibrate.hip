// Measurement aids. The VALU issue ceiling of the device: the trace kernels are bound by VALU
// instruction issue at partial lane occupancy, so their roofline is stated against the rate at
// which the SIMDs retire wave64 VALU instructions; k_valu_peak measures that rate with every lane
// active (eight independent FP32 FMA chains per lane, eight waves per SIMD). The L2 float-atomic
// ceiling: the grid adjoint is bound by its gradient atomics, so k_atomic_peak measures the rate of
// FP32 atomic adds without return to scattered words of eight per-XCD copies, the adjoint's layout.
#include "device.h"

namespace cupbr {
namespace {

constexpr int kCalBlock = 256;

// Each FMA reads one VGPR (its own chain), the scalar m and the inline constant 1.0, so no two
// vector operands of one instruction compete for a register bank.
__global__ __launch_bounds__(kCalBlock) void k_valu_peak(float* __restrict__ out, uint32_t iters, float m) {
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    float a0 = (float)gtid, a1 = a0 + 1.0f, a2 = a0 + 2.0f, a3 = a0 + 3.0f;
    float a4 = a0 + 4.0f, a5 = a0 + 5.0f, a6 = a0 + 6.0f, a7 = a0 + 7.0f;
    for (uint32_t k = 0; k < iters; ++k) {
        a0 = __builtin_fmaf(a0, m, 1.0f);
        a1 = __builtin_fmaf(a1, m, 1.0f);
        a2 = __builtin_fmaf(a2, m, 1.0f);
        a3 = __builtin_fmaf(a3, m, 1.0f);
        a4 = __builtin_fmaf(a4, m, 1.0f);
        a5 = __builtin_fmaf(a5, m, 1.0f);
        a6 = __builtin_fmaf(a6, m, 1.0f);
        a7 = __builtin_fmaf(a7, m, 1.0f);
    }
    out[gtid] = ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7));
}

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// block b adds into copy b % 8 (blocks reach the 8 XCDs round-robin); each lane's words are
// pseudo-random, so a wave's 64 atomics land on (nearly) 64 distinct cache lines
__global__ __launch_bounds__(kCalBlock) void k_atomic_peak(float* __restrict__ copies, uint32_t words, uint32_t iters) {
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    float* copy = copies + (size_t)(blockIdx.x & 7u) * words;
    for (uint32_t k = 0; k < iters; ++k) atomicAdd(copy + mix32(gtid * 0x9E3779B9u + k) % words, 1.0f);
}

}  // namespace
}  // namespace cupbr

using namespace cupbr;

extern "C" int cupbr_measure_atomic_peak(uint32_t iterations, uint32_t words_per_copy, double* milliseconds,
                                         uint64_t* atomics) {
    if (!milliseconds || !atomics || iterations == 0 || words_per_copy == 0 || words_per_copy > (1u << 26))
        return set_error(CUPBR_ERR_INVALID_ARGUMENT, "null output, zero iterations or words_per_copy outside 1..2^26");
    const int blocks = device_cu_count() * 8;
    DevBuf copies;
    if (int rc = copies.ensure((size_t)8 * words_per_copy * sizeof(float))) return rc;
    CUPBR_HIP_CHECK(hipMemset(copies.as<float>(), 0, (size_t)8 * words_per_copy * sizeof(float)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    CUPBR_HIP_CHECK(hipEventCreate(&e0));
    CUPBR_HIP_CHECK(hipEventCreate(&e1));
    hipLaunchKernelGGL(k_atomic_peak, dim3(blocks), dim3(kCalBlock), 0, nullptr, copies.as<float>(), words_per_copy,
                       iterations);
    CUPBR_HIP_CHECK(hipEventRecord(e0, nullptr));
    hipLaunchKernelGGL(k_atomic_peak, dim3(blocks), dim3(kCalBlock), 0, nullptr, copies.as<float>(), words_per_copy,
                       iterations);
    CUPBR_HIP_CHECK(hipEventRecord(e1, nullptr));
    CUPBR_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CUPBR_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *milliseconds = ms;
    *atomics = (uint64_t)blocks * kCalBlock * iterations;  // lane atomics of the timed launch
    return CUPBR_OK;
}

extern "C" int cupbr_measure_valu_peak(uint32_t iterations, double* milliseconds, uint64_t* fma_instructions) {
    if (!milliseconds || !fma_instructions || iterations == 0)
        return set_error(CUPBR_ERR_INVALID_ARGUMENT, "null output or zero iterations");
    const int blocks = device_cu_count() * 8;  // 32 waves per CU: 8 per SIMD
    const size_t n = (size_t)blocks * kCalBlock;
    DevBuf out;
    if (int rc = out.ensure(n * sizeof(float))) return rc;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    CUPBR_HIP_CHECK(hipEventCreate(&e0));
    CUPBR_HIP_CHECK(hipEventCreate(&e1));
    // one untimed launch, then the timed one
    hipLaunchKernelGGL(k_valu_peak, dim3(blocks), dim3(kCalBlock), 0, nullptr, out.as<float>(), iterations, 0.999f);
    CUPBR_HIP_CHECK(hipEventRecord(e0, nullptr));
    hipLaunchKernelGGL(k_valu_peak, dim3(blocks), dim3(kCalBlock), 0, nullptr, out.as<float>(), iterations, 0.999f);
    CUPBR_HIP_CHECK(hipEventRecord(e1, nullptr));
    CUPBR_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    CUPBR_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *milliseconds = ms;
    *fma_instructions = (uint64_t)(n / 64) * iterations * 8;  // wave64 FMA instructions of the timed launch
    return CUPBR_OK;
}
