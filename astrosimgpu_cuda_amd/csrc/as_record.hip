// as_record.hip -- spike recording: the spikes of held steps as an ordered
// list of (step, neuron) events, compacted on the device.
//
// The reference keeps only a spike total (src/astrosim_fused.cu:584); a
// recorder needs every spike. The engines hold the spike flags of their most
// recent steps as bit words (the plain method one 64-bit word per 64 neurons
// and step, the blocked method one 32-bit word per neuron and block, bit k =
// step k of the block). For a window of steps two kernels turn them into
// events, ordered by step, then neuron:
//   count_kernel: one workgroup per step sums the popcounts of its words;
//   write_kernel: one workgroup per step starts at the sum of the earlier
//     steps' counts and writes its neurons in order (a workgroup scan of the
//     words' popcounts gives each word its place).
// Both need device memory beyond the caller's outputs: the per-step counts
// and the events (the outputs are host arrays). as_sim_read_spikes allocates
// them per call; as_sim_read_spikes_noalloc stages them in `scratch`, a
// fixed device array, and allocates nothing.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "as_engine.h"

namespace astrosim {
namespace {

constexpr int kThreads = 256;

// Events the allocation-free path stages per window chunk: 2^20 (8 MB).
constexpr std::uint64_t kNoAllocEvents = std::uint64_t(1) << 20;

struct ScratchLayout {
    std::uint32_t counts[kRecordMaxSteps];
    std::uint32_t step_rel[kNoAllocEvents];
    std::uint32_t neuron[kNoAllocEvents];
};

// The allocation-free path's staging, one copy per GPU for the life of the
// process. Host code must not use this name as a device address: its device
// address comes from hipGetSymbolAddress (scratch_on_current_device).
static __device__ __align__(256) unsigned char scratch[sizeof(ScratchLayout)];

__device__ __forceinline__ unsigned long long step_word(const SpikeWords& w, int s, std::uint32_t word) {
    if (w.kind == 0) return static_cast<const unsigned long long*>(w.base)[w.off[s] + word];
    const std::uint32_t* p = static_cast<const std::uint32_t*>(w.base) + w.off[s];
    const std::uint32_t j0 = word * 64u;
    const std::uint32_t n = w.n_neurons - j0 < 64u ? w.n_neurons - j0 : 64u;
    unsigned long long m = 0ull;
    for (std::uint32_t k = 0; k < n; ++k) m |= static_cast<unsigned long long>((p[j0 + k] >> w.bit[s]) & 1u) << k;
    return m;
}

__device__ __forceinline__ std::uint32_t n_words(const SpikeWords& w) { return (w.n_neurons + 63u) / 64u; }

__global__ __launch_bounds__(kThreads) void count_kernel(SpikeWords w, std::uint32_t* counts) {
    __shared__ std::uint32_t total;
    if (threadIdx.x == 0) total = 0;
    __syncthreads();
    const int s = blockIdx.x;
    std::uint32_t c = 0;
    for (std::uint32_t i = threadIdx.x; i < n_words(w); i += blockDim.x) c += __popcll(step_word(w, s, i));
    for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o);
    if ((threadIdx.x & 63u) == 0 && c) atomicAdd(&total, c);
    __syncthreads();
    if (threadIdx.x == 0) counts[s] = total;
}

// Exclusive prefix sum of v over the workgroup; *sum receives the total.
__device__ std::uint32_t block_exclusive_scan(std::uint32_t v, std::uint32_t* sum) {
    __shared__ std::uint32_t wave_tot[kThreads / 64];
    const unsigned lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    std::uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const std::uint32_t y = __shfl_up(x, o);
        if (lane >= unsigned(o)) x += y;
    }
    if (lane == 63u) wave_tot[wave] = x;
    __syncthreads();
    std::uint32_t before = 0, all = 0;
    for (unsigned k = 0; k < blockDim.x / 64u; ++k) {
        if (k < wave) before += wave_tot[k];
        all += wave_tot[k];
    }
    __syncthreads();
    *sum = all;
    return before + x - v;
}

__global__ __launch_bounds__(kThreads) void write_kernel(SpikeWords w, const std::uint32_t* counts, std::uint64_t limit,
                                                         std::uint32_t* step_rel, std::uint32_t* neuron) {
    const int s = blockIdx.x;
    std::uint64_t base = 0;
    for (int k = 0; k < s; ++k) base += counts[k];
    const std::uint32_t nw = n_words(w);
    for (std::uint32_t c0 = 0; c0 < nw; c0 += blockDim.x) {
        const std::uint32_t i = c0 + threadIdx.x;
        unsigned long long m = i < nw ? step_word(w, s, i) : 0ull;
        std::uint32_t chunk = 0;
        std::uint64_t k = base + block_exclusive_scan(static_cast<std::uint32_t>(__popcll(m)), &chunk);
        while (m) {
            const unsigned b = __ffsll(static_cast<long long>(m)) - 1;
            m &= m - 1ull;
            if (k < limit) {
                step_rel[k] = static_cast<std::uint32_t>(s);
                neuron[k] = i * 64u + b;
            }
            ++k;
        }
        base += chunk;
    }
}

// Device address of this GPU's `scratch`, looked up once per device (with
// that device current) and kept.
void* scratch_on_current_device() {
    static std::mutex mu;
    static std::vector<void*> ptr;
    int dev = 0;
    AS_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    if (ptr.size() <= std::size_t(dev)) ptr.resize(std::size_t(dev) + 1, nullptr);
    if (!ptr[dev]) {
        void* p = nullptr;
        AS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptr[dev] = p;
    }
    return ptr[dev];
}

}  // namespace

void read_spikes(Engine& eng, std::int64_t first, std::int64_t n_steps, std::int64_t* steps, std::uint32_t* neurons,
                 std::uint64_t cap, std::uint64_t* count, bool no_alloc) {
    if (n_steps < 0 || first < 0) throw Error(AS_ERR_RANGE, "negative step window");
    if (no_alloc && cap > kNoAllocEvents)
        throw Error(AS_ERR_INVALID, "cap " + std::to_string(cap) + " exceeds the " + std::to_string(kNoAllocEvents) +
                                        " events the allocation-free path stages per call");
    if (cap > 0 && (!steps || !neurons)) throw Error(AS_ERR_INVALID, "NULL output arrays");
    hipStream_t s = eng.stream();
    std::uint32_t* counts = nullptr;
    std::uint32_t* ev_step = nullptr;
    std::uint32_t* ev_neuron = nullptr;
    if (no_alloc) {
        ScratchLayout* sc = static_cast<ScratchLayout*>(scratch_on_current_device());
        counts = sc->counts;
        ev_step = sc->step_rel;
        ev_neuron = sc->neuron;
    }
    std::vector<void*> owned;
    auto alloc = [&](std::size_t bytes) {
        void* p = nullptr;
        if (hipMalloc(&p, std::max<std::size_t>(bytes, 4)) != hipSuccess) {
            (void)hipGetLastError();
            throw Error(AS_ERR_NOMEM, "device staging of " + std::to_string(bytes) + " bytes for spike events");
        }
        owned.push_back(p);
        return p;
    };
    std::uint64_t total = 0, written = 0;
    try {
        if (!no_alloc) counts = static_cast<std::uint32_t*>(alloc(kRecordMaxSteps * sizeof(std::uint32_t)));
        for (std::int64_t c0 = 0; c0 < n_steps; c0 += kRecordMaxSteps) {
            const int n = static_cast<int>(std::min<std::int64_t>(kRecordMaxSteps, n_steps - c0));
            const SpikeWords w = eng.spike_words(first + c0, n);
            hipLaunchKernelGGL(count_kernel, dim3(n), dim3(kThreads), 0, s, w, counts);
            AS_HIP_CHECK(hipGetLastError());
            std::uint32_t hc[kRecordMaxSteps];
            AS_HIP_CHECK(hipMemcpyAsync(hc, counts, n * sizeof(std::uint32_t), hipMemcpyDeviceToHost, s));
            AS_HIP_CHECK(hipStreamSynchronize(s));
            std::uint64_t chunk = 0;
            for (int i = 0; i < n; ++i) chunk += hc[i];
            total += chunk;
            const std::uint64_t want = std::min<std::uint64_t>(chunk, cap - written);
            if (want == 0) continue;
            if (!no_alloc) {
                ev_step = static_cast<std::uint32_t*>(alloc(want * sizeof(std::uint32_t)));
                ev_neuron = static_cast<std::uint32_t*>(alloc(want * sizeof(std::uint32_t)));
            }
            hipLaunchKernelGGL(write_kernel, dim3(n), dim3(kThreads), 0, s, w, counts, want, ev_step, ev_neuron);
            AS_HIP_CHECK(hipGetLastError());
            // the steps' 32-bit offsets land in the first half of their int64
            // outputs and are widened in place, last first (no host buffer)
            std::int64_t* out = steps + written;
            std::uint32_t* rel = reinterpret_cast<std::uint32_t*>(out);
            AS_HIP_CHECK(hipMemcpyAsync(rel, ev_step, want * sizeof(std::uint32_t), hipMemcpyDeviceToHost, s));
            AS_HIP_CHECK(hipMemcpyAsync(neurons + written, ev_neuron, want * sizeof(std::uint32_t),
                                        hipMemcpyDeviceToHost, s));
            AS_HIP_CHECK(hipStreamSynchronize(s));
            for (std::uint64_t i = want; i-- > 0;) {
                std::uint32_t r;
                std::memcpy(&r, rel + i, sizeof r);
                const std::int64_t v = first + c0 + r;
                std::memcpy(out + i, &v, sizeof v);
            }
            written += want;
            if (!no_alloc) {
                for (void* p : {static_cast<void*>(ev_step), static_cast<void*>(ev_neuron)}) {
                    (void)hipFree(p);
                    owned.erase(std::find(owned.begin(), owned.end(), p));
                }
            }
        }
    } catch (...) {
        for (void* p : owned) (void)hipFree(p);
        throw;
    }
    for (void* p : owned) (void)hipFree(p);
    *count = total;
}

}  // namespace astrosim
