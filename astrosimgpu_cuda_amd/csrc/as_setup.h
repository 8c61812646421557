// as_setup.h -- host-side network construction: sequential setup RNG,
// randomised per-cell parameters, tripartite connectivity and its inversion
// into per-target CSR.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "as_model.h"

namespace astrosim {

// Stateful sequential draws on one (seed, stream), with a cached Box-Muller
// spare. Uses the host C library's log/sqrt/sin/cos exactly as the reference
// setup does. Reference: src/astrosim_fused.cu:180-208.
class HostRng {
public:
    HostRng(std::uint64_t seed, std::uint64_t stream) : seed_(seed), stream_(stream) {}
    double uniform() { return rng_uniform(seed_, stream_, counter_++); }
    double normal();
    // Advances the stream exactly as `count` normal() calls would.
    void skip_normals(std::uint64_t count);
    double normal_redraw(double mean, double stddev, double lo, double hi, int max_tries = 1000);
    std::uint64_t counter() const { return counter_; }

private:
    std::uint64_t seed_, stream_, counter_ = 0;
    bool has_spare_ = false;
    double spare_ = 0.0;
};

// mean * N(1, var) restricted to [lower, upper] * mean (bounds swapped for a
// negative mean). Reference: src/astrosim_fused.cu:210-216.
double draw_scaled(double mean, bool enabled, double lower, double upper, double var, HostRng& rng);

// Gap to the next Bernoulli(p) success among `remaining` candidates.
// Reference: src/astrosim_fused.cu:617-624. `log1m_p` = log1p(-p).
index_t geometric_skip(double p, double log1m_p, index_t remaining, HostRng& rng);

// An allocator whose resize() leaves trivial elements uninitialised: large
// arrays that are about to be overwritten are then neither zero-filled nor
// first touched by one thread.
template <class T>
struct NoInitAllocator : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInitAllocator<U>;
    };
    NoInitAllocator() = default;
    template <class U>
    NoInitAllocator(const NoInitAllocator<U>&) noexcept {}
    template <class U>
    void construct(U* p) noexcept {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... Args>
    void construct(U* p, Args&&... args) {
        ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
    }
};

// Connectivity inverted by target: row t lists the sources of target t in
// generation order. Reference: src/astrosim_fused.cu:296-315.
struct HostCSR {
    std::vector<index_t> row_start;                       // n_targets + 1
    std::vector<index_t, NoInitAllocator<index_t>> src;   // row_start[n_targets]
    index_t n_targets() const { return row_start.empty() ? 0 : index_t(row_start.size() - 1); }
    index_t n_edges() const { return index_t(src.size()); }
};

// A CSR in the memory of the GPU that generated it.
struct DeviceCSR {
    const index_t* row_start = nullptr;  // n_targets + 1
    const index_t* src = nullptr;        // n_edges
    index_t n_targets = 0;
    std::uint64_t n_edges = 0;
};

// Everything a simulation needs that depends only on (N_total, seed).
struct Network {
    ScalingConfig cfg;
    std::uint64_t seed = 0;
    HostCSR inc_exc;  // per neuron: excitatory sources (global ids < n_exc)
    HostCSR inc_inh;  // per neuron: inhibitory sources (LOCAL ids in [0, n_inh))
    HostCSR inc_sic;  // per neuron: astrocyte sources
    HostCSR inc_e2a;  // per astrocyte: excitatory neuron sources
    std::vector<double> Ca_tot, IP3_0, tau_IP3, delta_IP3;  // per astrocyte
    std::vector<double> V_reset, b;                          // per neuron
    double build_seconds = 0.0;  // connectivity generation + inversion
    double phase_seconds[4] = {0, 0, 0, 0};  // skip+pools, edge generation, inversion, parameters
    // Connectivity kept on the GPU by build_network_gpu (keep_on_device):
    // the four CSRs (exc, inh, sic, e2a) in device memory, freed with the
    // last copy of dev_owner; the host CSRs above are then empty. Only a
    // batch on the same GPU consumes such a network.
    DeviceCSR dev[4];
    std::shared_ptr<void> dev_owner;
    // A copy of the host CSRs in pinned host memory, held by pin_owner (a
    // server stages host-built networks so): uploads copy from it
    // asynchronously instead of from the vectors above.
    const index_t* pin_rs[4] = {nullptr, nullptr, nullptr, nullptr};
    const index_t* pin_src[4] = {nullptr, nullptr, nullptr, nullptr};
    std::shared_ptr<void> pin_owner;
    bool on_device() const { return dev_owner != nullptr; }
    bool host_csrs() const { return !on_device() || !inc_exc.row_start.empty(); }
    // Edges of CSR `which` (0 exc, 1 inh, 2 sic, 3 e2a), wherever it is held.
    std::uint64_t edges(int which) const {
        if (on_device()) return dev[which].n_edges;
        const HostCSR* c[4] = {&inc_exc, &inc_inh, &inc_sic, &inc_e2a};
        return c[which]->n_edges();
    }
};

// Builds the network for (n_total, seed). Throws std::invalid_argument for
// sizes the 32-bit edge ids cannot hold.
void build_network(index_t n_total, std::uint64_t seed, Network& net);

// The same network with the connectivity generated on the current GPU
// (as_gpubuild.hip); identical bit for bit. Returns false, with the reason
// in *why and `net` untouched, for networks that method declines (too small
// for its draw statistics, or a rare draw pattern it does not resolve); the
// caller then uses build_network.
// Smallest N_total for which the automatic choice (AS_BUILD_GPU) takes the
// GPU generator. Its latency is about constant (0.08-0.11 s from 20,000 to
// 100,000 cells on one MI355X), the host walk's grows with the network
// (0.07 s at 20,000, 0.21 s at 100,000), and alone they cross near 25,000.
// GPU builds share the device with running simulations, though: serving
// with the threshold at 30,000 measured slower than at 100,000.
constexpr index_t kGpuBuildAutoMin = 100000;

// `stream`: the HIP stream of the build's device work (nullptr: a stream of
// its own). `free_stream`: the stream the kept CSRs are freed on when the
// network goes (nullptr: `stream`; both nullptr: the null stream). A
// caller-given stream must outlive the network.
namespace gb { struct BuildStats; }
bool build_network_gpu(index_t n_total, std::uint64_t seed, Network& net, std::string* why,
                       gb::BuildStats* stats = nullptr, bool keep_on_device = false, hipStream_t stream = nullptr,
                       hipStream_t free_stream = nullptr);

// Moves a host-built network's CSRs into device memory of the current GPU
// (as build_network_gpu's keep_on_device leaves them; the host CSRs are then
// emptied). Leaves the network as it was when device memory is short.
// `stream` and `free_stream` as for build_network_gpu.
void stage_csrs_on_device(Network& net, hipStream_t stream = nullptr, hipStream_t free_stream = nullptr);

// Fills the host CSRs of a network whose connectivity build_network_gpu
// kept on the device (a copy; the device CSRs stay). No-op otherwise.
void materialize_host_csrs(Network& net);

// Astrocyte pools drawn from `rng` ([neuron][min(5, n_astro)] astrocyte ids).
// Reference: src/astrosim_fused.cu:651-671.
std::vector<index_t> make_pools(const ScalingConfig& cfg, HostRng& rng);

// Per-cell parameters (Ca_tot, IP3_0, tau_IP3, delta_IP3, V_reset, b) from a
// fresh setup stream. Reference: src/astrosim_fused.cu:917-933.
void draw_cell_params(const ScalingConfig& cfg, std::uint64_t seed, Network& net);

// Writes the edge list as CSV (type,source,target), astrocytes prefixed 'A',
// inhibitory sources shifted to global ids. Reference: src/astrosim_fused.cu:873-902.
bool dump_edges_csv(const Network& net, const std::string& path);

// Largest N_total accepted: keeps every CSR's edge count far below 2^32.
constexpr index_t kMaxNTotal = 40000000u;

}  // namespace astrosim
