// as_batch.h -- the blocked, event-driven engine over a batch of networks.
//
// A batch owns one device arena with fixed capacities (neurons, astrocytes,
// edges per connection type, members). Each member network occupies
// contiguous ranges of those spaces; ids are rebased to global ones, and a
// per-workgroup tile table maps each workgroup to one member. Members may be
// added and removed at block boundaries while others run (continuous
// batching); results of a member are bit-identical to running it alone.
#pragma once

#include <cstdint>
#include <vector>
#include <memory>

#include "as_engine.h"

namespace astrosim {

struct BatchCapacity {
    std::uint64_t neurons = 0, astros = 0;
    std::uint64_t edges[4] = {0, 0, 0, 0};  // exc, inh, sic, e2a
    std::uint32_t members = 1;
};

// Exact capacity for the given networks side by side.
BatchCapacity capacity_for(const Network* const* nets, int n);

class Batch {
public:
    virtual ~Batch() = default;
    // Whether `net` fits in the free ranges now.
    virtual bool fits(const Network& net) const = 0;
    // Admits `net` to run `n_steps` steps starting at the current global
    // step, which must be a block boundary. Returns the member slot.
    virtual int add(const Network& net, std::int64_t n_steps) = 0;
    // add for n networks at once: their uploads are enqueued back to back
    // and waited for once. members[i] receives network i's slot. If one
    // does not fit, the ones before it are admitted, the rest get -1 and
    // the error is thrown.
    virtual void add_many(const Network* const* nets, int n, const std::int64_t* n_steps, int* members) = 0;
    // Reserves `net`'s ranges and slot and uploads it on the batch's upload
    // stream, after the work enqueued so far and beside later steps; the
    // member joins at the first block boundary after the upload completes
    // (activate_ready), then runs `n_steps` steps. `net` must outlive the
    // activation. Returns the member slot.
    virtual int add_async(const Network& net, std::int64_t n_steps) = 0;
    // At a block boundary: activates every member whose upload has completed
    // and appends its slot to *out. Returns how many.
    virtual int activate_ready(std::vector<int>* out) = 0;
    // Retires a member; its ranges are reused once no in-flight block can
    // still deliver its emissions.
    virtual void remove(int member) = 0;
    // Advances every active member k steps (any k; partial blocks allowed).
    virtual void enqueue_steps(std::int64_t k) = 0;
    // Advances whole blocks through a captured hipGraph (global step must be
    // a block boundary). The graph is captured once per membership and block
    // count and replayed afterwards.
    virtual void run_blocks_graph(std::int64_t blocks) = 0;
    // Times candidate workgroup sizes on the next blocks of the live run and
    // keeps the fastest (results do not depend on it). Returns the choice.
    virtual int autotune(const int* candidates, int n, int blocks_each) = 0;
    virtual int block_threads() const = 0;
    // Steps one integrate launch advances (the LDS chunk in use).
    virtual int chunk_steps() const = 0;

    virtual std::uint64_t spike_count(int member) = 0;
    virtual void reset_ever_active(int member) = 0;
    virtual void read_state(int member, int field, double* out) = 0;
    virtual void read_refractory(int member, std::int32_t* out) = 0;
    virtual void read_ever_active(int member, std::uint8_t* out) = 0;
    virtual void read_spike_counts(int member, std::uint32_t* out) = 0;  // spikes per neuron
    virtual void read_raster(int member, std::int64_t first_local, int n, std::uint8_t* out) = 0;
    virtual SpikeWords spike_words(int member, std::int64_t first_local, int n) = 0;
    virtual std::int64_t member_steps(int member) const = 0;  // local steps enqueued
    // Local steps the member has run once the device reaches global step g.
    virtual std::int64_t member_steps_at(int member, std::int64_t g) const = 0;
    // A finished member's spike total and final calcium (hi + lo), read on
    // the batch's read stream: later blocks do not touch a member past its
    // last step, so the read runs beside them. The caller has waited for the
    // work that ran the member's last step.
    virtual void read_final(int member, std::uint64_t* spikes, double* ca) = 0;
    // Enqueues the copies read_final makes on the batch's block stream, after
    // the work enqueued so far (which must include the member's last step),
    // into pinned staging; the member's read_final then waits for these
    // copies only. A block stream needs no stream of its own for reads.
    virtual void prefetch_final(int member) = 0;
    virtual std::int64_t member_target(int member) const = 0;
    virtual bool member_active(int member) const = 0;
    virtual int active_members() const = 0;
    virtual std::uint64_t member_overflows() = 0;  // bucket overflows seen (diagnostic)

    virtual std::int64_t global_step() const = 0;
    virtual int block_len() const = 0;
    virtual hipStream_t stream() const = 0;
    virtual std::uint64_t device_bytes() const = 0;
    virtual std::uint64_t free_neurons() const = 0;
    virtual void set_timing(bool on) = 0;
    virtual void kernel_time(double* total_ms, std::int64_t* launches) = 0;
};

std::uint64_t batch_footprint(const BatchCapacity& cap, const RunParams& rp);
// s: the stream of the batch's blocks. io: the stream of member uploads and
// of result reads that were not prefetched (nullptr: two streams of the
// batch's own); a caller-given io stream may be shared by several batches
// and must outlive this one.
std::unique_ptr<Batch> make_batch(const RunParams& rp, const BatchCapacity& cap, hipStream_t s,
                                  hipStream_t io = nullptr);

}  // namespace astrosim
