// as_server.cpp -- continuous-batching simulation server.
//
// Requests (N_total, seed, steps, precision) arrive at any time. A pool of
// host threads builds each request's network (connectivity and parameters);
// a driver thread admits built networks into one of two resident batches
// (FP32c and FP64, each on its own HIP stream, capacities sized from an HBM
// budget) as soon as their ranges fit, advances both batches in rounds of
// blocks (the two streams overlap on the device; rounds are pipelined two
// deep), and retires members that reached their step count, reading their
// results and freeing their ranges for the next admissions. No request waits
// for a batch to drain.
//
// Streams and hardware queues: work on two HIP streams that share a hardware
// queue runs in submission order, so a build on a queue shared with a
// batch's blocks would wait for each round. Plain HIP streams share the
// process's GPU_MAX_HW_QUEUES queues (default 4) with every other stream the
// process has created, the least-used queue first, so which streams meet
// depends on what ran before the server. The server therefore creates its
// streams with a CU mask covering every CU: HIP gives such a stream a
// hardware queue of its own. Then each batch has a block stream, one upload
// stream serves both batches, and up to 8 build streams are each held by one
// builder at a time. If a masked stream cannot be created, the server falls
// back to plain streams and creates no more of them than the process has
// queues: one block stream per batch from 5 queues on, else one for both
// (the batches' rounds then run one after the other), one upload stream, and
// the rest as build streams (at 4 queues: 1 + 1 + 2).
// A builder that finds no build stream free builds a network below
// kGpuWaitMin cells with the host walk instead, and waits for a stream
// otherwise (the host walk takes about 2 s at 1M cells). Host-built networks
// need no stream: their CSRs are copied into pinned host memory (a pool the
// server keeps), and the upload copies them from there asynchronously.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/astrosim.h"
#include "as_batch.h"
#include "as_engine.h"
#include "as_setup.h"

namespace astrosim {
namespace {

using Clock = std::chrono::steady_clock;

// Networks from this size on wait for a free build stream rather than take
// the host walk (0.8 s at 400,000 cells against about 0.1 s on the GPU).
constexpr index_t kGpuWaitMin = 400000;

// Pinned host buffers, reused: a buffer returns to the pool when the network
// staged in it is dropped (after its upload has completed).
class PinnedPool {
public:
    ~PinnedPool() {
        for (auto& b : all_) (void)hipHostFree(b.first);
    }
    // A buffer of at least `bytes` (its size in *got).
    void* take(std::size_t bytes, std::size_t* got) {
        std::lock_guard<std::mutex> g(mu_);
        std::size_t best = free_.size();
        for (std::size_t i = 0; i < free_.size(); ++i)
            if (free_[i].second >= bytes && (best == free_.size() || free_[i].second < free_[best].second)) best = i;
        if (best < free_.size()) {
            auto b = free_[best];
            free_.erase(free_.begin() + std::ptrdiff_t(best));
            *got = b.second;
            return b.first;
        }
        const std::size_t sz = (std::max<std::size_t>(bytes, 1) + (std::size_t(4) << 20) - 1) & ~((std::size_t(4) << 20) - 1);
        void* p = nullptr;
        AS_HIP_CHECK(hipHostMalloc(&p, sz, hipHostMallocDefault));
        all_.emplace_back(p, sz);
        *got = sz;
        return p;
    }
    void give(void* p, std::size_t bytes) {
        std::lock_guard<std::mutex> g(mu_);
        free_.emplace_back(p, bytes);
    }

private:
    std::mutex mu_;
    std::vector<std::pair<void*, std::size_t>> all_, free_;
};

// Copies a host-built network's CSRs into one pinned buffer from `pool`.
void stage_pinned(Network& net, PinnedPool& pool) {
    const HostCSR* c[4] = {&net.inc_exc, &net.inc_inh, &net.inc_sic, &net.inc_e2a};
    std::size_t total = 0;
    for (const HostCSR* h : c) total += (h->row_start.size() + h->src.size()) * sizeof(index_t);
    std::size_t got = 0;
    char* p = static_cast<char*>(pool.take(total, &got));
    std::size_t off = 0;
    for (int k = 0; k < 4; ++k) {
        index_t* rs = reinterpret_cast<index_t*>(p + off);
        std::memcpy(rs, c[k]->row_start.data(), c[k]->row_start.size() * sizeof(index_t));
        off += c[k]->row_start.size() * sizeof(index_t);
        index_t* src = reinterpret_cast<index_t*>(p + off);
        if (!c[k]->src.empty()) std::memcpy(src, c[k]->src.data(), c[k]->src.size() * sizeof(index_t));
        off += c[k]->src.size() * sizeof(index_t);
        net.pin_rs[k] = rs;
        net.pin_src[k] = src;
    }
    PinnedPool* pp = &pool;
    net.pin_owner = std::shared_ptr<void>(p, [pp, got](void* q) { pp->give(q, got); });
}

// A stream on a hardware queue of its own (a full CU mask), or nullptr when
// HIP refuses one.
hipStream_t dedicated_stream() {
    int dev = 0;
    hipDeviceProp_t prop{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess ||
        prop.multiProcessorCount <= 0) {
        (void)hipGetLastError();
        return nullptr;
    }
    const int cus = prop.multiProcessorCount;
    std::vector<std::uint32_t> mask(std::size_t((cus + 31) / 32), 0xFFFFFFFFu);
    if (cus % 32 != 0) mask.back() = (1u << (cus % 32)) - 1u;
    hipStream_t s = nullptr;
    if (hipExtStreamCreateWithCUMask(&s, static_cast<std::uint32_t>(mask.size()), mask.data()) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return s;
}

// The process's hardware queues per device: GPU_MAX_HW_QUEUES, else HIP's
// default of 4.
int hw_queues() {
    const char* v = std::getenv("GPU_MAX_HW_QUEUES");
    const int q = v ? std::atoi(v) : 0;
    return q > 0 ? q : 4;
}

struct Req {
    as_request in{};
    as_request_result out{};
    std::unique_ptr<Network> net;
    int member = -1;
    int which = 0;  // batch: 0 FP32c, 1 FP64
};

// Capacity whose footprint fits `bytes`, with the scaling law's proportions
// (per neuron: 80 exc, 20 inh, 16 sic in-edges; per astrocyte 80 n2a) and 6%
// headroom for the binomial spread of edge counts.
BatchCapacity capacity_for_budget(std::uint64_t bytes, const RunParams& rp, std::uint32_t members) {
    auto cap_for = [&](std::uint64_t neurons) {
        BatchCapacity c;
        c.neurons = neurons;
        c.astros = neurons / 4 + 64;
        c.edges[0] = neurons * 85;
        c.edges[1] = neurons * 22;
        c.edges[2] = neurons * 18;
        c.edges[3] = neurons * 18;
        c.members = members;
        return c;
    };
    // the largest edge space (85 per neuron) must stay below 2^32 ids
    std::uint64_t lo = 1024, hi = 4000000000ull / 85;
    if (batch_footprint(cap_for(lo), rp) > bytes) throw Error(AS_ERR_NOMEM, "HBM budget too small for a batch");
    while (hi - lo > 1024) {
        const std::uint64_t mid = (lo + hi) / 2;
        if (batch_footprint(cap_for(mid), rp) <= bytes) lo = mid;
        else hi = mid;
    }
    return cap_for(lo);
}

}  // namespace

class Server {
public:
    Server(const as_server_config& c) : cfg_(c), t0_(Clock::now()) {
        as_config base;
        as_config_default(&base);
        base.dt = c.dt;
        base.block = c.block;
        base.method = AS_METHOD_BLOCKED;
        as_config c32 = base, c64 = base;
        c32.precision = AS_FP32C;
        c64.precision = AS_FP64;
        rp_[0] = resolve_params(c32);
        rp_[1] = resolve_params(c64);
        size_t free_b = 0, total_b = 0;
        AS_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
        const double frac = c.hbm_fraction > 0.0 && c.hbm_fraction < 0.95 ? c.hbm_fraction : 0.5;
        const std::uint64_t budget = static_cast<std::uint64_t>(frac * double(free_b));
        const double share64 = std::min(0.9, std::max(0.05, c.fp64_share));
        const std::uint32_t members = std::max(8, c.max_members);
        cap_[0] = capacity_for_budget(std::uint64_t(budget * (1.0 - share64)), rp_[0], members);
        cap_[1] = capacity_for_budget(std::uint64_t(budget * share64), rp_[1], members);
        // One stream per hardware queue, block streams first: the batches'
        // blocks, with the reads of finished members' results enqueued right
        // after the round that ends them (Batch::prefetch_final); the member
        // uploads of both batches (they build each member's out-CSR: device
        // work beside the blocks); the build streams.
        const int n_ded = 3 + std::min(8, std::max(1, c.build_threads));
        std::vector<hipStream_t> ded;
        for (int i = 0; i < n_ded; ++i) {
            hipStream_t st = dedicated_stream();
            if (!st) break;
            ded.push_back(st);
        }
        if (int(ded.size()) == n_ded) {
            dedicated_queues_ = true;
            stream_[0] = ded[0];
            stream_[1] = ded[1];
            io_ = ded[2];
            build_.assign(ded.begin() + 3, ded.end());
        } else {
            for (hipStream_t st : ded) (void)hipStreamDestroy(st);
            const int queues = hw_queues();
            const int n_block = queues >= 5 ? 2 : 1;
            AS_HIP_CHECK(hipStreamCreateWithFlags(&stream_[0], hipStreamNonBlocking));
            if (n_block == 2) AS_HIP_CHECK(hipStreamCreateWithFlags(&stream_[1], hipStreamNonBlocking));
            else stream_[1] = stream_[0];
            AS_HIP_CHECK(hipStreamCreateWithFlags(&io_, hipStreamNonBlocking));
            const int n_build = std::min(8, std::max(1, queues - n_block - 1));
            build_.assign(std::size_t(n_build), nullptr);
            for (hipStream_t& b : build_) AS_HIP_CHECK(hipStreamCreateWithFlags(&b, hipStreamNonBlocking));
        }
        free_build_ = build_;
        for (int i = 0; i < 2; ++i) {
            for (int k = 0; k < 2; ++k) {
                AS_HIP_CHECK(hipEventCreate(&round0_[k][i]));
                AS_HIP_CHECK(hipEventCreate(&round1_[k][i]));
            }
            batch_[i] = make_batch(rp_[i], cap_[i], stream_[i], io_);
        }
        setup_where_ = c.setup_where;
        const int nb = std::max(1, c.build_threads);
        for (int i = 0; i < nb; ++i) builders_.emplace_back([this] { build_loop(); });
        driver_ = std::thread([this] { drive_loop(); });
    }

    ~Server() {
        {
            std::lock_guard<std::mutex> g(mu_);
            stop_ = true;
        }
        stopping_ = true;
        cv_build_.notify_all();
        cv_ready_.notify_all();
        cv_stream_.notify_all();
        for (auto& t : builders_) t.join();
        if (driver_.joinable()) driver_.join();
        // networks not yet resident free their device CSRs on io_
        for (auto& q : reqs_) q->net.reset();
        for (int i = 0; i < 2; ++i) {
            batch_[i].reset();
            for (int k = 0; k < 2; ++k) {
                if (round0_[k][i]) (void)hipEventDestroy(round0_[k][i]);
                if (round1_[k][i]) (void)hipEventDestroy(round1_[k][i]);
            }
        }
        std::vector<hipStream_t> own = build_;
        own.insert(own.end(), {stream_[0], io_});
        if (stream_[1] != stream_[0]) own.push_back(stream_[1]);
        for (hipStream_t s : own) {
            if (!s) continue;
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }

    std::int64_t submit(const as_request& r) {
        if (r.n_total == 0 || r.n_total > kMaxNTotal) throw Error(AS_ERR_INVALID, "n_total out of range");
        if (r.n_steps < 0) throw Error(AS_ERR_INVALID, "n_steps must be >= 0");
        if (r.precision != AS_FP32C && r.precision != AS_FP64) throw Error(AS_ERR_INVALID, "unknown precision");
        std::lock_guard<std::mutex> g(mu_);
        if (!error_.empty()) throw Error(AS_ERR_HIP, "server stopped after an error: " + error_);
        if (stop_) throw Error(AS_ERR_INVALID, "server is stopping");
        auto q = std::make_shared<Req>();
        q->in = r;
        q->out.id = static_cast<std::int64_t>(reqs_.size());
        q->out.status = AS_REQ_BUILDING;
        q->out.t_submit = now();
        q->which = r.precision == AS_FP64 ? 1 : 0;
        reqs_.push_back(q);
        build_q_.push_back(q);
        cv_build_.notify_one();
        return q->out.id;
    }

    as_request_result result(std::int64_t id) {
        std::lock_guard<std::mutex> g(mu_);
        if (id < 0 || id >= std::int64_t(reqs_.size())) throw Error(AS_ERR_INVALID, "no such request");
        return reqs_[id]->out;
    }

    // Waits for every request to finish; an error of the driver ends the
    // wait at once and is returned (every unfinished request was rejected).
    bool wait_all(double timeout_s) {
        std::unique_lock<std::mutex> g(mu_);
        const bool done = cv_done_.wait_for(g, std::chrono::duration<double>(timeout_s),
                                            [&] { return finished_ == reqs_.size() || !error_.empty(); });
        if (!error_.empty()) throw Error(AS_ERR_HIP, "server stopped after an error: " + error_);
        return done;
    }

    as_server_stats stats() {
        std::lock_guard<std::mutex> g(mu_);
        if (!error_.empty()) throw Error(AS_ERR_HIP, "server stopped after an error: " + error_);
        as_server_stats s{};
        s.submitted = reqs_.size();
        std::vector<double> lat;
        double first_submit = 1e300, last_done = 0.0;
        for (auto& r : reqs_) {
            first_submit = std::min(first_submit, r->out.t_submit);
            if (r->out.status == AS_REQ_DONE) {
                lat.push_back(r->out.t_done - r->out.t_submit);
                last_done = std::max(last_done, r->out.t_done);
            }
            if (r->out.status == AS_REQ_REJECTED) ++s.rejected;
        }
        s.completed = lat.size();
        if (!lat.empty()) {
            std::sort(lat.begin(), lat.end());
            auto pct = [&](double p) {
                const double k = p * (lat.size() - 1);
                const std::size_t i = static_cast<std::size_t>(k);
                const double f = k - i;
                return i + 1 < lat.size() ? lat[i] * (1 - f) + lat[i + 1] * f : lat[i];
            };
            s.p50_latency_s = pct(0.50);
            s.p99_latency_s = pct(0.99);
            s.mean_latency_s = 0;
            for (double v : lat) s.mean_latency_s += v;
            s.mean_latency_s /= lat.size();
            s.sims_per_s = last_done > first_submit ? lat.size() / (last_done - first_submit) : 0.0;
        }
        s.device_bytes = batch_[0]->device_bytes() + batch_[1]->device_bytes();
        s.peak_active = peak_active_;
        s.peak_resident_cells = peak_cells_;
        s.capacity_cells = cap_[0].neurons + cap_[0].astros + cap_[1].neurons + cap_[1].astros;
        s.gpu_busy_s = gpu_busy_s_;
        s.admit_s = admit_s_;
        s.retire_s = retire_s_;
        s.cell_steps = cell_steps_;
        s.dedicated_queues = dedicated_queues_ ? 1 : 0;
        s.build_streams = static_cast<std::int32_t>(build_.size());
        return s;
    }

private:
    double now() const { return std::chrono::duration<double>(Clock::now() - t0_).count(); }

    // A free build stream, or nullptr when none is free and `wait` is false
    // (also when the server stops).
    hipStream_t take_build_stream(bool wait) {
        std::unique_lock<std::mutex> g(stream_mu_);
        if (wait) cv_stream_.wait(g, [&] { return stopping_.load() || !free_build_.empty(); });
        if (free_build_.empty()) return nullptr;
        hipStream_t s = free_build_.back();
        free_build_.pop_back();
        return s;
    }

    void give_build_stream(hipStream_t s) {
        if (!s) return;
        {
            std::lock_guard<std::mutex> g(stream_mu_);
            free_build_.push_back(s);
        }
        cv_stream_.notify_one();
    }

    void build_loop() {
        for (;;) {
            std::shared_ptr<Req> q;
            {
                std::unique_lock<std::mutex> g(mu_);
                cv_build_.wait(g, [&] { return stop_ || !build_q_.empty(); });
                if (stop_) return;
                q = build_q_.front();
                build_q_.pop_front();
                q->out.t_build_start = now();
            }
            auto net = std::make_unique<Network>();
            bool ok = true;
            hipStream_t bs = nullptr;
            try {
                // a GPU-built network stays in device memory until its admission
                // copies it; its CSRs are freed on io_ (the driver drops it)
                const bool only_gpu = setup_where_ == AS_BUILD_GPU_ONLY;
                const bool try_gpu = only_gpu || (setup_where_ == AS_BUILD_GPU && q->in.n_total >= kGpuBuildAutoMin);
                if (try_gpu) bs = take_build_stream(only_gpu || q->in.n_total >= kGpuWaitMin);
                if (!bs || !build_network_gpu(q->in.n_total, q->in.seed, *net, nullptr, nullptr, true, bs, io_)) {
                    if (only_gpu) throw std::runtime_error(bs ? "GPU build declined" : "server stopping");
                    give_build_stream(bs);
                    bs = nullptr;
                    build_network(q->in.n_total, q->in.seed, *net);
                    stage_pinned(*net, pinned_);
                }
            } catch (...) {
                ok = false;
            }
            give_build_stream(bs);
            std::lock_guard<std::mutex> g(mu_);
            // the driver rejected it meanwhile (it stopped after an error)
            if (q->out.status != AS_REQ_BUILDING) continue;
            q->out.n_neurons = ok ? net->cfg.n_neurons() : 0;
            q->out.n_astro = ok ? net->cfg.n_astro : 0;
            q->out.built_on_gpu = ok && net->on_device() ? 1 : 0;
            q->out.t_built = now();
            if (!ok) {
                q->out.status = AS_REQ_REJECTED;
                ++finished_;
                cv_done_.notify_all();
                continue;
            }
            q->net = std::move(net);
            q->out.status = AS_REQ_QUEUED;
            ready_[q->which].push_back(q);
            cv_ready_.notify_one();
        }
    }

    // Admits queued requests in arrival order while they fit. A request that
    // can never fit the empty batch is rejected.
    void admit() {
        for (int b = 0; b < 2; ++b) {
            for (;;) {
                std::shared_ptr<Req> q;
                {
                    std::lock_guard<std::mutex> g(mu_);
                    if (ready_[b].empty()) break;
                    q = ready_[b].front();
                }
                Batch& B = *batch_[b];
                if (!B.fits(*q->net)) {
                    const Network* p = q->net.get();
                    const BatchCapacity need = capacity_for(&p, 1);
                    const bool never = need.neurons > cap_[b].neurons || need.astros > cap_[b].astros ||
                                       need.edges[0] > cap_[b].edges[0] || need.edges[1] > cap_[b].edges[1] ||
                                       need.edges[2] > cap_[b].edges[2] || need.edges[3] > cap_[b].edges[3];
                    if (!never) break;  // wait for ranges to free up (FIFO: no overtaking)
                    std::lock_guard<std::mutex> g(mu_);
                    ready_[b].pop_front();
                    q->out.status = AS_REQ_REJECTED;
                    q->net.reset();
                    ++finished_;
                    cv_done_.notify_all();
                    continue;
                }
                // uploaded beside the running steps; it joins once the upload is done
                q->member = B.add_async(*q->net, q->in.n_steps);
                std::lock_guard<std::mutex> g(mu_);
                ready_[b].pop_front();
                staging_[b].push_back(q);
            }
        }
    }

    // Requests whose upload has completed join their batch (at the current
    // block boundary) and start running.
    void activate() {
        for (int b = 0; b < 2; ++b) {
            if (staging_[b].empty()) continue;
            std::vector<int> slots;
            if (batch_[b]->activate_ready(&slots) == 0) continue;
            std::vector<std::shared_ptr<Req>> still;
            std::lock_guard<std::mutex> g(mu_);
            for (auto& q : staging_[b]) {
                if (std::find(slots.begin(), slots.end(), q->member) == slots.end()) {
                    still.push_back(q);
                    continue;
                }
                q->net.reset();  // resident in HBM now
                q->out.status = AS_REQ_RUNNING;
                q->out.t_admit = now();
                running_[b].push_back(q);
            }
            staging_[b].swap(still);
        }
    }

    // Retires the requests that have run all their steps by each batch's
    // last completed round.
    void retire() {
        for (int b = 0; b < 2; ++b) {
            Batch& B = *batch_[b];
            std::vector<std::shared_ptr<Req>> still;
            for (auto& q : running_[b]) {
                if (B.member_steps_at(q->member, done_step_[b]) < q->in.n_steps) {
                    still.push_back(q);
                    continue;
                }
                std::uint64_t spikes = 0;
                std::vector<double> ca(q->out.n_astro);
                B.read_final(q->member, &spikes, ca.data());
                double mean = 0.0;
                std::uint32_t active = 0;
                for (double c : ca) {
                    mean += c;
                    if ((c - rp_[b].SIC_th) * 1000.0 > 1.0) ++active;
                }
                B.remove(q->member);
                std::lock_guard<std::mutex> g(mu_);
                q->out.spikes = spikes;
                q->out.mean_ca = ca.empty() ? 0.0 : mean / ca.size();
                q->out.active_astro = active;
                q->out.status = AS_REQ_DONE;
                q->out.t_done = now();
                cell_steps_ += double(q->out.n_neurons + q->out.n_astro) * double(q->in.n_steps);
                ++finished_;
                cv_done_.notify_all();
            }
            running_[b].swap(still);
        }
    }

    // Rounds are pipelined two deep: round k+1 is enqueued before the driver
    // waits for round k, so the device runs k+1 while the host admits,
    // retires (the results of members that finished by the end of k were
    // copied right after k) and rebuilds the tile tables for k+2.
    void drive_loop() {
        const int blocks = std::max(1, cfg_.blocks_per_round);
        int cur = 0, prev = 0;
        bool inflight = false;  // round `prev` is enqueued and not yet completed
        try {
            for (;;) {
                {
                    std::unique_lock<std::mutex> g(mu_);
                    cv_ready_.wait_for(g, std::chrono::milliseconds(2), [&] {
                        return stop_ || inflight || !ready_[0].empty() || !ready_[1].empty() ||
                               !running_[0].empty() || !running_[1].empty() || !staging_[0].empty() ||
                               !staging_[1].empty();
                    });
                    if (stop_) return;
                }
                auto a0 = Clock::now();
                activate();
                admit_s_ += std::chrono::duration<double>(Clock::now() - a0).count();
                Round& R = rounds_[cur];
                R = Round();
                for (int b = 0; b < 2; ++b) {
                    if (running_[b].empty()) continue;
                    R.ran[b] = true;
                    for (auto& q : running_[b]) R.cells += q->out.n_neurons + q->out.n_astro;
                    R.active += int(running_[b].size());
                    AS_HIP_CHECK(hipEventRecord(round0_[cur][b], stream_[b]));
                    batch_[b]->enqueue_steps(std::int64_t(blocks) * batch_[b]->block_len());  // async
                    AS_HIP_CHECK(hipEventRecord(round1_[cur][b], stream_[b]));
                    R.end_step[b] = batch_[b]->global_step();
                    // the results of members this round finishes are read
                    // right after it on the same stream
                    for (auto& q : running_[b])
                        if (batch_[b]->member_steps_at(q->member, R.end_step[b]) >= q->in.n_steps)
                            batch_[b]->prefetch_final(q->member);
                }
                const bool now_ran = R.ran[0] || R.ran[1];
                // Admission's host work runs while the rounds execute; its
                // uploads are ordered after them (Batch::add_async) and the
                // members join at a later round's start.
                a0 = Clock::now();
                admit();
                admit_s_ += std::chrono::duration<double>(Clock::now() - a0).count();
                if (!inflight) {
                    if (now_ran) {
                        inflight = true;  // keep it queued and enqueue the next
                        prev = cur;
                        cur ^= 1;
                    } else if (!staging_[0].empty() || !staging_[1].empty()) {
                        // only uploads in flight: poll them without spinning
                        std::this_thread::sleep_for(std::chrono::microseconds(200));
                    }
                    continue;
                }
                if (cfg_.fail_after_rounds > 0 && ++rounds_done_ >= cfg_.fail_after_rounds)
                    throw Error(AS_ERR_HIP, "injected failure after " + std::to_string(rounds_done_) + " rounds");
                // complete round `prev`; the one just enqueued keeps the device busy
                const Round& P = rounds_[prev];
                int first = -1;
                for (int b = 0; b < 2; ++b) {
                    if (!P.ran[b]) continue;
                    AS_HIP_CHECK(hipEventSynchronize(round1_[prev][b]));
                    done_step_[b] = P.end_step[b];
                    if (first < 0) first = b;
                }
                // the round's span: from the first batch's start to the
                // later end of the two (on two streams they overlap)
                float round_ms = 0.0f;
                for (int b = 0; b < 2; ++b) {
                    if (!P.ran[b]) continue;
                    float ms = 0.0f;
                    AS_HIP_CHECK(hipEventElapsedTime(&ms, round0_[prev][first], round1_[prev][b]));
                    round_ms = std::max(round_ms, ms);
                }
                gpu_busy_s_ += 1e-3 * double(round_ms);
                peak_active_ = std::max<std::uint64_t>(peak_active_, P.active);
                peak_cells_ = std::max<std::uint64_t>(peak_cells_, P.cells);
                const auto r0 = Clock::now();
                retire();
                retire_s_ += std::chrono::duration<double>(Clock::now() - r0).count();
                if (now_ran) {
                    prev = cur;
                    cur ^= 1;
                } else {
                    inflight = false;
                }
            }
        } catch (const std::exception& e) {
            // let the device finish what was enqueued before stopping
            for (int b = 0; b < 2; ++b) (void)hipStreamSynchronize(stream_[b]);
            std::lock_guard<std::mutex> g(mu_);
            error_ = e.what();
            if (error_.empty()) error_ = "unknown error";
            stop_ = true;
            // fail every unfinished request so waiters return; builders skip
            // the requests they were building (their status is final now)
            for (auto& r : reqs_)
                if (r->out.status != AS_REQ_DONE && r->out.status != AS_REQ_REJECTED) {
                    r->out.status = AS_REQ_REJECTED;
                    ++finished_;
                }
            build_q_.clear();
            cv_done_.notify_all();
            cv_build_.notify_all();
        }
    }

    as_server_config cfg_;
    Clock::time_point t0_;
    RunParams rp_[2];
    BatchCapacity cap_[2];
    hipStream_t stream_[2] = {nullptr, nullptr};  // each batch's blocks and result reads (may be one stream)
    hipStream_t io_ = nullptr;                     // member uploads of both batches; frees of built CSRs
    std::vector<hipStream_t> build_;               // the builders' device work, one builder per stream
    std::vector<hipStream_t> free_build_;          // build streams no builder holds
    bool dedicated_queues_ = false;                // the streams have hardware queues of their own
    std::mutex stream_mu_;
    std::condition_variable cv_stream_;
    PinnedPool pinned_;  // host-built networks' CSRs until their upload
    std::atomic<bool> stopping_{false};
    // A round: up to `blocks_per_round` blocks of each batch with members.
    struct Round {
        bool ran[2] = {false, false};
        std::int64_t end_step[2] = {0, 0};  // each batch's global step at its end
        std::uint64_t cells = 0;
        int active = 0;
    };
    Round rounds_[2];
    hipEvent_t round0_[2][2] = {}, round1_[2][2] = {};  // [round slot][batch]: the round's span
    std::int64_t done_step_[2] = {0, 0};                // each batch's step at its last completed round
    std::unique_ptr<Batch> batch_[2];
    std::mutex mu_;
    std::condition_variable cv_build_, cv_ready_, cv_done_;
    std::vector<std::shared_ptr<Req>> reqs_;
    std::deque<std::shared_ptr<Req>> build_q_, ready_[2];
    std::vector<std::shared_ptr<Req>> staging_[2];  // admitted, upload in flight
    std::vector<std::shared_ptr<Req>> running_[2];
    std::size_t finished_ = 0;
    bool stop_ = false;
    std::string error_;
    int setup_where_ = AS_BUILD_GPU;
    std::vector<std::thread> builders_;
    std::thread driver_;
    std::uint64_t peak_active_ = 0, peak_cells_ = 0;
    int rounds_done_ = 0;
    double gpu_busy_s_ = 0.0, cell_steps_ = 0.0, admit_s_ = 0.0, retire_s_ = 0.0;
};

}  // namespace astrosim

// The C ABI of the server lives with the other entry points' conventions.
struct as_server {
    std::unique_ptr<astrosim::Server> s;
};

namespace {
template <class F>
as_status server_guard(F&& f) {
    try {
        f();
        return AS_OK;
    } catch (const astrosim::Error& e) {
        astrosim::set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        astrosim::set_last_error(e.what());
        return AS_ERR_INVALID;
    }
}
}  // namespace

extern "C" {

void as_server_config_default(as_server_config* c) {
    if (!c) return;
    c->dt = 0.1;
    c->block = 128;
    c->hbm_fraction = 0.5;
    c->fp64_share = 0.3;
    c->build_threads = 8;
    c->blocks_per_round = 16;
    c->max_members = 512;
    c->setup_where = AS_BUILD_GPU;
    c->fail_after_rounds = 0;
}

as_status as_server_create(const as_server_config* cfg, as_server** out) {
    return server_guard([&] {
        if (!cfg || !out) throw astrosim::Error(AS_ERR_INVALID, "NULL argument");
        *out = nullptr;
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
            (void)hipGetLastError();
            throw astrosim::Error(AS_ERR_NODEVICE, "No GPU device visible.");
        }
        std::unique_ptr<as_server> s(new as_server);
        s->s.reset(new astrosim::Server(*cfg));
        *out = s.release();
    });
}

as_status as_server_submit(as_server* s, const as_request* r, int64_t* id) {
    return server_guard([&] {
        if (!s || !r || !id) throw astrosim::Error(AS_ERR_INVALID, "NULL argument");
        *id = s->s->submit(*r);
    });
}

as_status as_server_result(as_server* s, int64_t id, as_request_result* out) {
    return server_guard([&] {
        if (!s || !out) throw astrosim::Error(AS_ERR_INVALID, "NULL argument");
        *out = s->s->result(id);
    });
}

as_status as_server_wait_all(as_server* s, double timeout_s, int32_t* all_done) {
    return server_guard([&] {
        if (!s || !all_done) throw astrosim::Error(AS_ERR_INVALID, "NULL argument");
        *all_done = s->s->wait_all(timeout_s) ? 1 : 0;
    });
}

as_status as_server_get_stats(as_server* s, as_server_stats* out) {
    return server_guard([&] {
        if (!s || !out) throw astrosim::Error(AS_ERR_INVALID, "NULL argument");
        *out = s->s->stats();
    });
}

void as_server_destroy(as_server* s) { delete s; }

}  // extern "C"
