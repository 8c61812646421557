// as_bw_probe.hip -- the HBM rates a streaming kernel reaches on the device it
// runs on (a measurement tool; the library does not use it): the ceiling the
// plain method's CSR streaming is compared with.
//
// A read-only kernel loads a 5 GiB buffer as 16-byte words (one load in
// flight per lane per iteration, or four) at several grid sizes and keeps
// the fastest of five runs of each; a device-to-device copy of the same
// buffer gives the read-plus-write rate.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

#define BW_CHECK(x)                                                                      \
    do {                                                                                 \
        const hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                          \
            std::printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 1;                                                                    \
        }                                                                                \
    } while (0)

// XOR of every word a lane reads; one store per lane keeps the loads.
__global__ void read_kernel(const uint4* __restrict__ p, std::size_t n, unsigned* out) {
    unsigned acc = 0;
    const std::size_t stride = std::size_t(gridDim.x) * blockDim.x;
    for (std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint4 v = p[i];
        acc ^= v.x ^ v.y ^ v.z ^ v.w;
    }
    out[std::size_t(blockIdx.x) * blockDim.x + threadIdx.x] = acc;
}

// The same with four loads in flight per lane.
__global__ void read4_kernel(const uint4* __restrict__ p, std::size_t n, unsigned* out) {
    unsigned acc = 0;
    const std::size_t stride = std::size_t(gridDim.x) * blockDim.x;
    std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const uint4 a = p[i], b = p[i + stride], c = p[i + 2 * stride], d = p[i + 3 * stride];
        acc ^= a.x ^ a.y ^ a.z ^ a.w ^ b.x ^ b.y ^ b.z ^ b.w ^ c.x ^ c.y ^ c.z ^ c.w ^ d.x ^ d.y ^ d.z ^ d.w;
    }
    for (; i < n; i += stride) {
        const uint4 a = p[i];
        acc ^= a.x ^ a.y ^ a.z ^ a.w;
    }
    out[std::size_t(blockIdx.x) * blockDim.x + threadIdx.x] = acc;
}

}  // namespace

int main() {
    const std::size_t bytes = std::size_t(5) << 30;
    const std::size_t n = bytes / sizeof(uint4);
    const int grids[] = {1024, 2048, 4096, 8192, 16384, 65536};
    uint4* p = nullptr;
    void* q = nullptr;
    unsigned* out = nullptr;
    BW_CHECK(hipMalloc(&p, bytes));
    BW_CHECK(hipMalloc(&q, bytes));
    BW_CHECK(hipMalloc(&out, std::size_t(65536) * 256 * sizeof(unsigned)));
    BW_CHECK(hipMemset(p, 1, bytes));
    hipEvent_t a, b;
    BW_CHECK(hipEventCreate(&a));
    BW_CHECK(hipEventCreate(&b));
    for (int k = 0; k < 2; ++k) {
        for (int g : grids) {
            float best = 1e30f;
            for (int r = 0; r < 5; ++r) {
                BW_CHECK(hipEventRecord(a));
                if (k == 0) hipLaunchKernelGGL(read_kernel, dim3(g), dim3(256), 0, 0, p, n, out);
                else hipLaunchKernelGGL(read4_kernel, dim3(g), dim3(256), 0, 0, p, n, out);
                BW_CHECK(hipGetLastError());
                BW_CHECK(hipEventRecord(b));
                BW_CHECK(hipEventSynchronize(b));
                float ms = 0.0f;
                BW_CHECK(hipEventElapsedTime(&ms, a, b));
                if (ms < best) best = ms;
            }
            std::printf("%-28s grid %6d x 256: %7.3f ms  %.2f TB/s read\n",
                        k ? "read, 4 loads per lane" : "read, 1 load per lane", g, best,
                        double(bytes) / (best * 1e-3) / 1e12);
        }
    }
    float best = 1e30f;
    for (int r = 0; r < 5; ++r) {
        BW_CHECK(hipEventRecord(a));
        BW_CHECK(hipMemcpyAsync(q, p, bytes, hipMemcpyDeviceToDevice, 0));
        BW_CHECK(hipEventRecord(b));
        BW_CHECK(hipEventSynchronize(b));
        float ms = 0.0f;
        BW_CHECK(hipEventElapsedTime(&ms, a, b));
        if (ms < best) best = ms;
    }
    std::printf("device copy of 5 GiB: %.3f ms  %.2f TB/s read + write\n", best,
                2.0 * double(bytes) / (best * 1e-3) / 1e12);
    BW_CHECK(hipFree(p));
    BW_CHECK(hipFree(q));
    BW_CHECK(hipFree(out));
    return 0;
}
