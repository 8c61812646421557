// as_selftest.hip -- test-only library (lib/libastrosim_selftest.so, not part
// of the C ABI in include/astrosim.h). It evaluates the deterministic math
// the simulator's FP64 path is built on -- exp, log and cos of as_dmath.h and
// the fixed-divisor division of as_model.h -- over arrays, once with the
// device code and once with the host code, so that tests can compare the two
// bit for bit over whole input ranges (subnormal and overflowing results,
// signed zeros, quotients next to rounding midpoints). It also evaluates the
// Gaussian draws: the FP64 one on both sides, and the FP32 one FP32 state
// uses (device only), for a bound on their difference.
#include <hip/hip_runtime.h>

#include <cstddef>

#include "as_engine.h"
#include "as_model.h"
#include "as_model32.h"

namespace {

enum Fn { kExp = 0, kLog = 1, kCos = 2, kDivide = 3, kDivideIeee = 4, kNormal = 5, kNormal32 = 6 };

__host__ __device__ inline double eval(int fn, double x, double y) {
    switch (fn) {
    case kExp: return as_exp(x);
    case kLog: return as_log(x);
    case kCos: return as_cos(x);
    case kDivide: return astrosim::divide(x, astrosim::make_fixed_divisor(y));
    case kDivideIeee: return x / y;
    default: return astrosim::rng_normal_z(as_dm_bits(x));
    }
}

__global__ void eval_kernel(int fn, const double* __restrict__ x, const double* __restrict__ y,
                            double* __restrict__ out, std::size_t n) {
    const std::size_t i = blockIdx.x * std::size_t(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    out[i] = fn == kNormal32 ? static_cast<double>(astrosim::normal32(as_dm_bits(x[i]))) : eval(fn, x[i], y[i]);
}

}  // namespace

// out[i] = f(x[i], y[i]) on the host. fn: 0 exp(x), 1 log(x), 2 cos(x),
// 3 x / y by the fixed-divisor method (y > 0), 4 x / y, 5 the FP64 Gaussian
// draw of the stream whose counter-0 finaliser input has the bits of x,
// 6 the FP32 one (device only). Returns 0, or -1 for an fn this side lacks.
extern "C" int as_selftest_eval_host(int fn, const double* x, const double* y, double* out, std::size_t n) {
    if (fn < kExp || fn > kNormal) return -1;
    for (std::size_t i = 0; i < n; ++i) out[i] = eval(fn, x[i], y[i]);
    return 0;
}

// The same on the current device (host arrays in and out). Returns 0, -1 for
// an unknown fn, or the HIP error code.
extern "C" int as_selftest_eval_device(int fn, const double* x, const double* y, double* out, std::size_t n) {
    if (fn < kExp || fn > kNormal32) return -1;
    if (n == 0) return 0;
    double *dx = nullptr, *dy = nullptr, *dout = nullptr;
    const std::size_t bytes = n * sizeof(double);
    hipError_t e = hipMalloc(&dx, bytes);
    if (e == hipSuccess) e = hipMalloc(&dy, bytes);
    if (e == hipSuccess) e = hipMalloc(&dout, bytes);
    if (e == hipSuccess) e = hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dy, y, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const unsigned threads = 256;
        const unsigned blocks = static_cast<unsigned>((n + threads - 1) / threads);
        hipLaunchKernelGGL(eval_kernel, dim3(blocks), dim3(threads), 0, 0, fn, dx, dy, dout, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(dx);
    (void)hipFree(dy);
    (void)hipFree(dout);
    return e == hipSuccess ? 0 : static_cast<int>(e);
}
