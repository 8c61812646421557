// as_engine.h -- internal interface between the C ABI and the GPU engines.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/astrosim.h"
#include "as_setup.h"

namespace astrosim {

// Error carrying an as_status code.
struct Error : std::runtime_error {
    as_status code;
    Error(as_status c, const std::string& msg) : std::runtime_error(msg), code(c) {}
};

// Message of the last failed C-ABI call on this thread (as_last_error).
void set_last_error(const std::string& msg);
const char* last_error();

[[noreturn]] void throw_hip(hipError_t e, const char* file, int line, const char* expr);

#define AS_HIP_CHECK(call)                                                        \
    do {                                                                          \
        hipError_t as_e_ = (call);                                                \
        if (as_e_ != hipSuccess) ::astrosim::throw_hip(as_e_, __FILE__, __LINE__, #call); \
    } while (0)

// Every constant a step needs, resolved once on the host from the run
// options and the biology defaults. Host-side precomputation uses the same
// as_dmath functions the oracle evaluates per cell, so values are identical.
struct RunParams {
    double dt = 0.1, h_step = 0.1, h_sixth = 0.1 / 6.0;
    int substeps = 1, block = 128;
    int precision = AS_FP64, method = AS_METHOD_PLAIN;
    double w_e = 5.0, w_i = -5.0, w_n2a = 0.2, w_a2n = 1.0;
    bool stp_enabled = true;
    double stp_U = 0.5, stp_tau_rec = 800.0, stp_tau_fac = 0.0;
    int d_e = 10, d_i = 10, d_a2n = 10, ring_slots = 11;
    int steps_per_noise = 10, sic_interval = 1;
    // Poisson drive: lambda = rate*dt*1e-3 and its Knuth limit exp(-lambda).
    double lam_astro = 0.0, lim_astro = 1.0, w_poiss_astro = 1.0;
    double lam_exc = 0.0, lim_exc = 1.0, w_poiss_exc = 1.0;
    double lam_inh = 0.0, lim_inh = 1.0, w_poiss_inh = 1.0;
    bool poiss_astro = false, poiss_exc = false, poiss_inh = false;
    double noise_astro = 0.0, noise_exc = 0.0, noise_inh = 0.0;
    // Alpha-cascade propagators exp(-h/tau) per population.
    double ex_decay_e = 0.0, in_decay_e = 0.0, ex_decay_i = 0.0, in_decay_i = 0.0;
    double SIC_th = 0.19669;
    AstroConst ac{};
    NeuronConst ne{}, ni{};
};

// Validates an as_config and resolves it. Throws Error(AS_ERR_INVALID).
RunParams resolve_params(const as_config& cfg);

// Bump layout of one device allocation (256-byte aligned pieces).
struct Layout {
    std::size_t total = 0;
    std::size_t add(std::size_t bytes) {
        const std::size_t off = total;
        total += (bytes + 255) & ~std::size_t(255);
        return off;
    }
};

// Where the spike flags of a window of held steps live on the device, for
// the recorder (as_record.hip). kind 0: one 64-bit word per 64 neurons and
// step (bit j & 63 of word j >> 6 at base + off[i]); kind 1: one 32-bit word
// per neuron and block (bit bit[i] of word j at base + off[i]).
constexpr int kRecordMaxSteps = 64;
struct SpikeWords {
    int kind = 0;
    const void* base = nullptr;
    std::uint32_t n_neurons = 0;
    int n_steps = 0;
    std::uint64_t off[kRecordMaxSteps] = {};
    std::uint32_t bit[kRecordMaxSteps] = {};
};

// One simulation on one HIP stream.
class Engine {
public:
    virtual ~Engine() = default;
    virtual void enqueue_steps(std::int64_t k) = 0;
    virtual std::uint64_t spike_count() = 0;  // synchronises
    virtual void reset_ever_active() = 0;
    virtual void read_state(int field, double* out) = 0;  // synchronises
    virtual void read_refractory(std::int32_t* out) = 0;
    virtual void read_ever_active(std::uint8_t* out) = 0;
    virtual void read_spike_counts(std::uint32_t* out) = 0;  // spikes per neuron since creation
    virtual void read_raster(std::int64_t first, int n, std::uint8_t* out) = 0;
    // Spike words of held steps [first, first + n), n <= kRecordMaxSteps
    // (AS_ERR_RANGE when not held).
    virtual SpikeWords spike_words(std::int64_t first, int n) = 0;
    virtual std::uint64_t device_bytes() const = 0;

    std::int64_t steps() const { return step_; }
    hipStream_t stream() const { return stream_; }
    index_t n_neurons() const { return n_neurons_; }
    index_t n_astro() const { return n_astro_; }

    // Kernel timing (HIP events around each dominant-kernel launch).
    virtual void set_timing(bool on);
    virtual void kernel_time(double* total_ms, std::int64_t* launches);

protected:
    // Brackets one launch of the engine's dominant step kernel with HIP
    // events on the engine's stream (no-ops while timing is off).
    void timing_begin();
    void timing_end(std::int64_t launches);

    hipStream_t stream_ = nullptr;
    std::int64_t step_ = 0;
    index_t n_neurons_ = 0, n_astro_ = 0, n_exc_ = 0;
    bool timing_ = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events_;
    hipEvent_t pending_start_ = nullptr;
    std::int64_t timed_launches_ = 0;
};

// The spikes of steps [first, first + n_steps) as (step, neuron) events in
// step, then neuron order: the first min(total, cap) into steps/neurons,
// the total into *count. no_alloc: staged in a fixed device array instead
// of per-call allocations (cap <= 2^20). Synchronises the engine's stream.
void read_spikes(Engine& eng, std::int64_t first, std::int64_t n_steps, std::int64_t* steps, std::uint32_t* neurons,
                 std::uint64_t cap, std::uint64_t* count, bool no_alloc);

std::uint64_t plain_footprint(const Network& net, const RunParams& rp);
std::unique_ptr<Engine> make_plain_engine(const Network& net, const RunParams& rp, hipStream_t s);
std::uint64_t blocked_footprint(const Network& net, const RunParams& rp);
std::unique_ptr<Engine> make_blocked_engine(const Network& net, const RunParams& rp, hipStream_t s);

}  // namespace astrosim
