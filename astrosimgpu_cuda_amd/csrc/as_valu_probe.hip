// as_valu_probe.hip -- throughput of the vector-instruction kinds the step
// kernels are made of, on the device it runs on (a measurement tool; the
// library does not use it).
//
// Each kernel runs 8 independent dependency chains per lane over a fixed
// number of iterations, 16 workgroups of 256 threads per CU (4 waves per
// SIMD), so the issue rate rather than a chain's latency sets the time.
// Printed: SIMD cycles per wave instruction = elapsed cycles x SIMDs /
// (wave-instructions executed), with the instructions per chain iteration
// of each kernel given below (as the compiler emits them for gfx950; check
// with --save-temps when changing a kernel).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

constexpr int kIters = 2048;
constexpr int kChains = 8;

#define PROBE_CHECK(x)                                                                   \
    do {                                                                                 \
        const hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                          \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 1;                                                                    \
        }                                                                                \
    } while (0)

// Op kinds (instructions per chain iteration):
//  0 v_fma_f32 (1)           1 v_fma_f64 (1)          2 v_add_f64 (1)
//  3 v_mul_f32 + v_med3_f32 (2)                       4 v_lshrrev_b32 + v_xad_u32 (2)
//  5 v_mul_hi_u32 (1)        6 v_exp_f32 (1)          7 v_add_f32 + v_rcp_f32 (2)
//  8 v_fma_f32 + v_fma_f64 (2, two chains interleaved)
template <int OP>
__global__ __launch_bounds__(256) void probe_kernel(std::uint32_t* out, std::uint32_t seed) {
    std::uint32_t a[kChains];
    double d[kChains];
    float f[kChains];
    for (int c = 0; c < kChains; ++c) {
        a[c] = (threadIdx.x * (2u * c + 3u)) ^ seed;
        d[c] = a[c] * 1e-9;
        f[c] = a[c] * 1e-9f;
    }
    const std::uint32_t k = seed | 1u;
    for (int i = 0; i < kIters; ++i) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) {
            if constexpr (OP == 0) f[c] = __builtin_fmaf(f[c], 0.999f, 0.25f);
            else if constexpr (OP == 1) d[c] = __builtin_fma(d[c], 0.999, 0.25);
            else if constexpr (OP == 2) d[c] = d[c] + 0.25;
            else if constexpr (OP == 3) f[c] = __builtin_amdgcn_fmed3f(f[c] * 1.0001f, -1.0f, 1.0f);
            else if constexpr (OP == 4) a[c] = (a[c] ^ (a[c] >> 3)) + k;
            else if constexpr (OP == 5) a[c] = __umulhi(a[c] | 1u, k + a[c]);
            else if constexpr (OP == 6) f[c] = __builtin_amdgcn_exp2f(f[c]);
            else if constexpr (OP == 7) f[c] = __builtin_amdgcn_rcpf(f[c] + 1.5f);
            else if constexpr (OP == 8) {
                f[c] = __builtin_fmaf(f[c], 0.999f, 0.25f);
                d[c] = __builtin_fma(d[c], 0.999, 0.25);
            }
        }
    }
    std::uint32_t r = 0;
    for (int c = 0; c < kChains; ++c) r ^= a[c] ^ std::uint32_t(d[c] * 1e6) ^ std::uint32_t(f[c] * 1e6f);
    out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

template <int OP>
int run(const char* name, int insts_per_iter, std::uint32_t* out, int blocks, double clock_hz, int cus) {
    hipEvent_t e0, e1;
    PROBE_CHECK(hipEventCreate(&e0));
    PROBE_CHECK(hipEventCreate(&e1));
    hipLaunchKernelGGL(probe_kernel<OP>, dim3(blocks), dim3(256), 0, 0, out, 1u);  // warm-up
    PROBE_CHECK(hipGetLastError());
    constexpr int kReps = 5;
    PROBE_CHECK(hipEventRecord(e0));
    for (int r = 0; r < kReps; ++r) hipLaunchKernelGGL(probe_kernel<OP>, dim3(blocks), dim3(256), 0, 0, out, 1u);
    PROBE_CHECK(hipEventRecord(e1));
    PROBE_CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    PROBE_CHECK(hipEventElapsedTime(&ms, e0, e1));
    const double wave_insts = double(blocks) * (256 / 64) * kReps * kIters * kChains * insts_per_iter;
    const double simd_cycles = ms * 1e-3 * clock_hz * cus * 4;
    std::printf("%-28s %8.3f ms  %5.2f SIMD cycles per wave instruction\n", name, ms, simd_cycles / wave_insts);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return 0;
}

}  // namespace

int main() {
    hipDeviceProp_t p;
    PROBE_CHECK(hipGetDeviceProperties(&p, 0));
    const int cus = p.multiProcessorCount;
    const double clock_hz = double(p.clockRate) * 1e3;
    std::printf("%s: %d CUs, %.0f MHz; 8 chains per lane, 4 waves per SIMD\n", p.gcnArchName, cus, clock_hz / 1e6);
    const int blocks = cus * 16;
    std::uint32_t* out = nullptr;
    PROBE_CHECK(hipMalloc(&out, std::size_t(blocks) * 256 * sizeof(std::uint32_t)));
    int rc = 0;
    rc |= run<0>("v_fma_f32", 1, out, blocks, clock_hz, cus);
    rc |= run<1>("v_fma_f64", 1, out, blocks, clock_hz, cus);
    rc |= run<2>("v_add_f64", 1, out, blocks, clock_hz, cus);
    rc |= run<3>("v_mul_f32 + v_med3_f32", 2, out, blocks, clock_hz, cus);
    rc |= run<4>("v_lshrrev_b32 + v_xad_u32", 2, out, blocks, clock_hz, cus);
    rc |= run<5>("v_mul_hi_u32 (+ v_or, v_add)", 3, out, blocks, clock_hz, cus);
    rc |= run<6>("v_exp_f32", 1, out, blocks, clock_hz, cus);
    rc |= run<7>("v_add_f32 + v_rcp_f32", 2, out, blocks, clock_hz, cus);
    rc |= run<8>("v_fma_f32 + v_fma_f64", 2, out, blocks, clock_hz, cus);
    PROBE_CHECK(hipFree(out));
    return rc;
}
