// as_gpubuild.hip -- the parallel connectivity generator (as_gpubuild.h,
// as_gpubuild_run.h) on the current GPU: the decode, chain, path, repair and
// row walks as kernels of one lane per item; hipCUB for the exclusive scans
// and the stable radix sorts that invert the generation-ordered edges into
// target-major rows (src/astrosim_fused.cu:307-314). Device memory comes from
// the stream-ordered allocator on a stream of the build's own, so builds on
// several host threads neither share state nor wait for other streams'
// work (a plain hipFree would wait for the whole device).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <chrono>
#include <future>
#include <stdexcept>
#include <string>
#include <thread>

#include "as_gpubuild_run.h"
#include "as_setup.h"

namespace astrosim {
namespace {

using namespace gb;

template <class F>
__global__ __launch_bounds__(256) void for_each_kernel(F f, u64 n) {
    const u64 i = u64(blockIdx.x) * 256u + threadIdx.x;
    if (i < n) f(i);
}

struct HipFailure {
    std::string what;
};

#define GB_CHECK(expr)                                                                            \
    do {                                                                                          \
        const hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                                     \
            throw HipFailure{std::string(#expr) + ": " + hipGetErrorString(e_) + " (" __FILE__ ":" + \
                             std::to_string(__LINE__) + ")"};                                     \
    } while (0)

class HipBackend {
public:
    // on the caller's stream, or on one of its own (nullptr)
    explicit HipBackend(hipStream_t s) : s_(s), own_(s == nullptr) {
        if (own_) GB_CHECK(hipStreamCreateWithFlags(&s_, hipStreamNonBlocking));
    }
    ~HipBackend() {
        (void)hipStreamSynchronize(s_);
        if (own_) (void)hipStreamDestroy(s_);
    }
    // A failed allocation declines the build (the caller falls back to the
    // host walk) rather than failing it.
    template <class T>
    T* alloc(u64 n) {
        void* p = nullptr;
        if (hipMallocAsync(&p, n * sizeof(T), s_) != hipSuccess) {
            (void)hipGetLastError();
            throw gb::Declined{"device memory exhausted (" + std::to_string(n * sizeof(T)) + " bytes)"};
        }
        return static_cast<T*>(p);
    }
    void release(void* p) { (void)hipFreeAsync(p, s_); }
    void h2d(void* d, const void* h, u64 bytes) {
        if (bytes == 0) return;
        GB_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s_));
        GB_CHECK(hipStreamSynchronize(s_));
    }
    void d2h(void* h, const void* d, u64 bytes) {
        if (bytes == 0) return;
        GB_CHECK(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, s_));
        GB_CHECK(hipStreamSynchronize(s_));
    }
    void zero(void* d, u64 bytes) {
        if (bytes) GB_CHECK(hipMemsetAsync(d, 0, bytes, s_));
    }
    template <class F>
    void run(u64 n, const F& f) {
        if (n == 0) return;
        const u64 blocks = (n + 255) / 256;
        if (blocks >= (u64(1) << 31)) throw HipFailure{"launch too large"};
        hipLaunchKernelGGL(for_each_kernel<F>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s_, f, n);
        GB_CHECK(hipGetLastError());
    }
    void scan_u64(u64* p, u64 n) { scan(p, n); }
    void scan_u32(u32* p, u64 n) { scan(p, n); }
    void sort_pairs(const u32* kin, const u32* vin, u32* kout, u32* vout, u64 n, int bits) {
        size_t bytes = 0;
        GB_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, static_cast<int>(n), 0,
                                                    bits, s_));
        void* tmp = alloc<char>(bytes + 16);
        GB_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, bytes, kin, kout, vin, vout, static_cast<int>(n), 0, bits,
                                                    s_));
        release(tmp);
    }

private:
    template <class T>
    void scan(T* p, u64 n) {
        if (n >= (u64(1) << 31)) throw HipFailure{"scan too large"};
        size_t bytes = 0;
        GB_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, p, p, static_cast<int>(n), s_));
        void* tmp = alloc<char>(bytes + 16);
        GB_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, bytes, p, p, static_cast<int>(n), s_));
        release(tmp);
    }
    hipStream_t s_ = nullptr;
    bool own_ = true;
};

}  // namespace

bool build_network_gpu(index_t n_total, std::uint64_t seed, Network& net, std::string* why,
                       gb::BuildStats* stats, bool keep_on_device, hipStream_t stream, hipStream_t free_stream) {
    if (!free_stream) free_stream = stream;
    std::string reason;
    auto decline = [&](const std::string& r) {
        if (why) *why = r;
        return false;
    };
    if (n_total == 0 || n_total > kMaxNTotal) throw std::invalid_argument("N_total out of range");
    const ScalingConfig cfg = make_scaling(n_total);
    // the method needs geometric skips (p < 1) and the rejection-sampled pools
    if (cfg.n_exc <= 80 || cfg.n_astro <= 4 * ScalingConfig::kPoolSize) return decline("network too small");
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0) return decline("no GPU");

    const auto t0 = std::chrono::steady_clock::now();
    Network out;
    out.cfg = cfg;
    out.seed = seed;
    // the per-cell parameters come from their own stream: drawn meanwhile
    gb::BuildStats st;
    std::thread params([&] {
        const auto tp = std::chrono::steady_clock::now();
        draw_cell_params(cfg, seed, out);
        st.params_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - tp).count();
    });
    bool ok = false;
    try {
        HostRng rng(seed, kSetupStream);
        rng.skip_normals(4ull * cfg.n_astro + 2ull * cfg.n_neurons());
        // the pools take at least kPoolSize draws per neuron: the edge draws
        // start no earlier than that, and the GPU decodes from there while
        // the pools are drawn
        const std::uint64_t c0_lo = rng.counter() + std::uint64_t(ScalingConfig::kPoolSize) * cfg.n_neurons();
        auto pools = std::async(std::launch::async, [&cfg, rng, t0, &out, &st]() mutable {
            std::vector<index_t> pool = make_pools(cfg, rng);
            out.phase_seconds[0] = st.pools_seconds =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            return std::make_pair(rng.counter(), std::move(pool));
        });
        auto pools_fn = [&pools](std::vector<index_t>& pool) {
            auto r = pools.get();
            pool = std::move(r.second);
            return r.first;
        };
        std::vector<void*> kept;
        {
            HipBackend bk(stream);  // its destructor waits for the build's stream
            ok = gb::build_connectivity(bk, cfg, seed, c0_lo, pools_fn, ScalingConfig::kPoolSize, out, st, reason,
                                        keep_on_device ? &kept : nullptr);
        }
        if (pools.valid()) pools.wait();
        if (ok && keep_on_device) {
            auto* held = new std::vector<void*>(std::move(kept));
            out.dev_owner = std::shared_ptr<void>(held, [free_stream](void* p) {
                auto* v = static_cast<std::vector<void*>*>(p);
                for (void* q : *v) (void)hipFreeAsync(q, free_stream);
                delete v;
            });
        }
    } catch (const HipFailure& f) {
        params.join();
        throw std::runtime_error("GPU network build: " + f.what);
    } catch (...) {
        params.join();
        throw;
    }
    const auto t_join = std::chrono::steady_clock::now();
    params.join();
    st.params_wait_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_join).count();
    if (stats) *stats = st;
    if (!ok) return decline(reason);
    const auto t1 = std::chrono::steady_clock::now();
    out.build_seconds = std::chrono::duration<double>(t1 - t0).count();
    out.phase_seconds[1] = st.seconds;
    net = std::move(out);
    return true;
}

void stage_csrs_on_device(Network& net, hipStream_t stream, hipStream_t free_stream) {
    if (net.on_device()) return;
    if (!free_stream) free_stream = stream;
    HostCSR* c[4] = {&net.inc_exc, &net.inc_inh, &net.inc_sic, &net.inc_e2a};
    hipStream_t s = stream;
    if (!s && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        return;  // left on the host
    }
    auto* held = new std::vector<void*>();
    bool ok = true;
    for (int k = 0; k < 4 && ok; ++k) {
        void* rs = nullptr;
        void* src = nullptr;
        const std::size_t nr = c[k]->row_start.size(), ne = c[k]->src.size();
        ok = hipMallocAsync(&rs, nr * sizeof(index_t), s) == hipSuccess &&
             hipMallocAsync(&src, std::max<std::size_t>(ne, 1) * sizeof(index_t), s) == hipSuccess;
        if (rs) held->push_back(rs);
        if (src) held->push_back(src);
        if (!ok) break;
        ok = hipMemcpyAsync(rs, c[k]->row_start.data(), nr * sizeof(index_t), hipMemcpyHostToDevice, s) ==
                 hipSuccess &&
             (ne == 0 ||
              hipMemcpyAsync(src, c[k]->src.data(), ne * sizeof(index_t), hipMemcpyHostToDevice, s) == hipSuccess);
        net.dev[k] = DeviceCSR{static_cast<const index_t*>(rs), static_cast<const index_t*>(src),
                               static_cast<index_t>(nr - 1), ne};
    }
    ok = hipStreamSynchronize(s) == hipSuccess && ok;
    if (!ok) {
        (void)hipGetLastError();
        for (void* p : *held) (void)hipFreeAsync(p, s);
        (void)hipStreamSynchronize(s);
        delete held;
        for (auto& d : net.dev) d = DeviceCSR{};
    } else {
        net.dev_owner = std::shared_ptr<void>(held, [free_stream](void* p) {
            auto* v = static_cast<std::vector<void*>*>(p);
            for (void* q : *v) (void)hipFreeAsync(q, free_stream);
            delete v;
        });
        for (HostCSR* h : c) *h = HostCSR();
    }
    if (!stream) (void)hipStreamDestroy(s);
}

void materialize_host_csrs(Network& net) {
    if (net.host_csrs()) return;
    HostCSR* c[4] = {&net.inc_exc, &net.inc_inh, &net.inc_sic, &net.inc_e2a};
    for (int k = 0; k < 4; ++k) {
        const DeviceCSR& d = net.dev[k];
        c[k]->row_start.resize(std::size_t(d.n_targets) + 1);
        c[k]->src.resize(d.n_edges);
        hipError_t e = hipMemcpy(c[k]->row_start.data(), d.row_start, c[k]->row_start.size() * sizeof(index_t),
                                 hipMemcpyDeviceToHost);
        if (e == hipSuccess && d.n_edges)
            e = hipMemcpy(c[k]->src.data(), d.src, d.n_edges * sizeof(index_t), hipMemcpyDeviceToHost);
        if (e != hipSuccess)
            throw std::runtime_error(std::string("copying a GPU-built network to the host: ") + hipGetErrorString(e));
    }
}

}  // namespace astrosim
