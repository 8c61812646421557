// as_selftest.hip -- test-only library (lib/libastrosim_selftest.so, not part
// of the C ABI in include/astrosim.h). It evaluates the deterministic math
// the simulator's FP64 path is built on -- exp, log and cos of as_dmath.h and
// the fixed-divisor division of as_model.h -- over arrays, once with the
// device code and once with the host code, so that tests can compare the two
// bit for bit over whole input ranges (subnormal and overflowing results,
// signed zeros, quotients next to rounding midpoints). It also evaluates the
// Gaussian draws: the FP64 one on both sides, and the FP32 one FP32 state
// uses (device only), for a bound on their difference; and the FP32 state's
// AdEx spike current (device) against the same current in long double
// (host), for its error in ulps. And it runs the
// parallel connectivity generator of as_gpubuild_run.h over a host backend
// (loops for kernels, std::stable_sort for the radix sort), so that its
// method is checked against the sequential builder without a GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "as_engine.h"
#include "as_gpubuild_run.h"
#include "as_model.h"
#include "as_model32.h"
#include "as_setup.h"

namespace {

enum Fn { kExp = 0, kLog = 1, kCos = 2, kDivide = 3, kDivideIeee = 4, kNormal = 5, kNormal32 = 6, kSpike32 = 7,
          kSpikeExact = 8 };

// The default configuration's neuron constants of population y (0
// excitatory, else inhibitory).
const astrosim::NeuronConst& neuron_const(const astrosim::RunParams& rp, double y) { return y == 0.0 ? rp.ne : rp.ni; }

astrosim::RunParams default_params() {
    as_config c;
    as_config_default(&c);
    c.precision = AS_FP32C;
    c.method = AS_METHOD_BLOCKED;
    return astrosim::resolve_params(c);
}

__global__ void spike32_kernel(astrosim::NeuronConst32 ne, astrosim::NeuronConst32 ni, const double* __restrict__ x,
                               const double* __restrict__ y, double* __restrict__ out, std::size_t n) {
    const std::size_t i = blockIdx.x * std::size_t(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const astrosim::NeuronConst32& p = y[i] == 0.0 ? ne : ni;
    out[i] = static_cast<double>(astrosim::spike_current32(p, static_cast<float>(x[i])));
}

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
// 6 the FP32 one (device only), 7 the FP32 state's spike current
// g_L*Delta_T/C_m*exp((V - V_T)/Delta_T) at V = float(x) for population y
// (device only), 8 the same in long double (host only). Returns 0, or -1
// for an fn this side lacks.
extern "C" int as_selftest_eval_host(int fn, const double* x, const double* y, double* out, std::size_t n) {
    if (fn == kSpikeExact) {
        const astrosim::RunParams rp = default_params();
        for (std::size_t i = 0; i < n; ++i) {
            const astrosim::NeuronConst& p = neuron_const(rp, y[i]);
            const long double V = static_cast<float>(x[i]);
            out[i] = static_cast<double>((long double)p.g_L * p.Delta_T / p.C_m *
                                         expl((V - (long double)p.V_th) / (long double)p.Delta_T));
        }
        return 0;
    }
    if (fn < kExp || fn > kNormal) return -1;
    for (std::size_t i = 0; i < n; ++i) out[i] = eval(fn, x[i], y[i]);
    return 0;
}

// The same on the current device (host arrays in and out). Returns 0, -1 for
// an unknown fn, or the HIP error code.
extern "C" int as_selftest_eval_device(int fn, const double* x, const double* y, double* out, std::size_t n) {
    if (fn < kExp || fn > kSpike32) return -1;
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
        if (fn == kSpike32) {
            const astrosim::Params32 q = astrosim::make_params32(default_params());
            hipLaunchKernelGGL(spike32_kernel, dim3(blocks), dim3(threads), 0, 0, q.ne, q.ni, dx, dy, dout, n);
        } else {
            hipLaunchKernelGGL(eval_kernel, dim3(blocks), dim3(threads), 0, 0, fn, dx, dy, dout, n);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost);
    (void)hipFree(dx);
    (void)hipFree(dy);
    (void)hipFree(dout);
    return e == hipSuccess ? 0 : static_cast<int>(e);
}

namespace {

using namespace astrosim::gb;

// Host stand-in for the GPU backend: every "kernel" is a loop over its items
// (split over a few threads; the bodies use atomics where items share data).
class HostBackend {
public:
    template <class T>
    T* alloc(u64 n) {
        return static_cast<T*>(::operator new(n * sizeof(T)));
    }
    void release(void* p) { ::operator delete(p); }
    void h2d(void* d, const void* h, u64 bytes) { if (bytes) std::memcpy(d, h, bytes); }
    void d2h(void* h, const void* d, u64 bytes) { if (bytes) std::memcpy(h, d, bytes); }
    void zero(void* d, u64 bytes) { if (bytes) std::memset(d, 0, bytes); }
    template <class F>
    void run(u64 n, const F& f) {
        const u64 T = n < 4096 ? 1 : 8;
        std::vector<std::thread> th;
        for (u64 t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                for (u64 i = n * t / T; i < n * (t + 1) / T; ++i) f(i);
            });
        for (auto& x : th) x.join();
    }
    void scan_u64(u64* p, u64 n) { scan(p, n); }
    void scan_u32(u32* p, u64 n) { scan(p, n); }
    void sort_pairs(const u32* kin, const u32* vin, u32* kout, u32* vout, u64 n, int) {
        std::vector<u64> idx(n);
        std::iota(idx.begin(), idx.end(), u64(0));
        std::stable_sort(idx.begin(), idx.end(), [&](u64 a, u64 b) { return kin[a] < kin[b]; });
        for (u64 i = 0; i < n; ++i) {
            kout[i] = kin[idx[i]];
            vout[i] = vin[idx[i]];
        }
    }

private:
    template <class T>
    void scan(T* p, u64 n) {
        T run = 0;
        for (u64 i = 0; i < n; ++i) {
            const T v = p[i];
            p[i] = run;
            run += v;
        }
    }
};

bool same_csr(const astrosim::HostCSR& a, const astrosim::HostCSR& b) {
    return a.row_start == b.row_start && a.src.size() == b.src.size() &&
           std::equal(a.src.begin(), a.src.end(), b.src.begin());
}

}  // namespace

// Builds the network (n_total, seed) with the sequential host builder and
// again with the parallel generator -- on the host backend (where = 0) or on
// the current GPU through build_network_gpu (where = 1) -- and compares the
// four CSRs (and, for where = 1, the per-cell parameters). Returns 0 when
// identical, 1 when different, 2 when the parallel generator declined, -1
// on an exception. stats (32 doubles, may be NULL): counters, chains,
// ambiguous skips, host rows of passes 0-2, records of passes 0 and 2,
// repairs of passes 0 and 2, largest index shift, generator seconds, host
// builder seconds, edges compared, mask of what differs (1 exc, 2 inh, 4 sic,
// 8 e2a, 16 parameters), then the generator's phase marks (decode, chains,
// passes 0-2, counts, emission, four inversions; seconds from its start),
// the pools' seconds and the wait for the parameter draws. cap_b (0: the
// default) is the number of row starts a chain stores past its region on
// the host backend; a small one makes most chains walk the rest again.
extern "C" int as_selftest_build_compare(uint32_t n_total, uint64_t seed, int where, double* stats,
                                         uint32_t cap_b) {
    using namespace astrosim;
    try {
        Network ref;
        const auto t0 = std::chrono::steady_clock::now();
        build_network(n_total, seed, ref);
        const double t_host = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        Network got;
        BuildStats st;
        std::string why;
        bool ok;
        if (where == 0) {
            const ScalingConfig cfg = make_scaling(n_total);
            got.cfg = cfg;
            got.seed = seed;
            HostRng rng(seed, kSetupStream);
            rng.skip_normals(4ull * cfg.n_astro + 2ull * cfg.n_neurons());
            const u32 take = static_cast<u32>(std::min<index_t>(ScalingConfig::kPoolSize, cfg.n_astro));
            const u64 c0_lo = rng.counter() + u64(take) * cfg.n_neurons();
            auto pools_fn = [&](std::vector<index_t>& pool) {
                pool = make_pools(cfg, rng);
                return rng.counter();
            };
            HostBackend bk;
            ok = build_connectivity(bk, cfg, seed, c0_lo, pools_fn, take, got, st, why, nullptr,
                                    cap_b == 0 ? kCapB : cap_b);
        } else {
            // as the library's own GPU builds: connectivity kept in device memory
            ok = build_network_gpu(n_total, seed, got, &why, &st, true);
            st.seconds = got.build_seconds;
            if (ok) materialize_host_csrs(got);
        }
        if (stats) {
            for (int k = 0; k < 11; ++k) stats[16 + k] = st.mark[k];
            stats[27] = st.pools_seconds;
            stats[28] = st.params_wait_seconds;
            stats[29] = double(st.max_b_draws);
            stats[30] = double(st.max_repair_rows);
            stats[31] = st.params_seconds;
            const double v[16] = {double(st.counters), double(st.chains), double(st.ambiguous),
                                  double(st.host_rows[0]), double(st.host_rows[1]), double(st.host_rows[2]),
                                  double(st.records[0]), double(st.records[2]), double(st.repairs[0]),
                                  double(st.repairs[2]), double(st.max_shift), st.seconds, t_host,
                                  double(ref.inc_exc.n_edges() + ref.inc_inh.n_edges() + ref.inc_sic.n_edges() +
                                         ref.inc_e2a.n_edges()),
                                  0.0, 0.0};
            std::memcpy(stats, v, sizeof v);
        }
        if (!ok) return 2;
        int diff = (same_csr(ref.inc_exc, got.inc_exc) ? 0 : 1) | (same_csr(ref.inc_inh, got.inc_inh) ? 0 : 2) |
                   (same_csr(ref.inc_sic, got.inc_sic) ? 0 : 4) | (same_csr(ref.inc_e2a, got.inc_e2a) ? 0 : 8);
        if (where == 1 && !(ref.Ca_tot == got.Ca_tot && ref.IP3_0 == got.IP3_0 && ref.tau_IP3 == got.tau_IP3 &&
                            ref.delta_IP3 == got.delta_IP3 && ref.V_reset == got.V_reset && ref.b == got.b))
            diff |= 16;
        if (stats) stats[14] = diff;
        return diff == 0 ? 0 : 1;
    } catch (...) {
        return -1;
    }
}
