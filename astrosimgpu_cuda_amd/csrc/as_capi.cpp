// as_capi.cpp -- the extern "C" boundary (include/astrosim.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <string>

#include "../../include/astrosim.h"
#include "as_batch.h"
#include "as_engine.h"
#include "as_setup.h"

struct as_network {
    astrosim::Network net;
    bool built_on_gpu = false;
    std::mutex mu;  // host CSRs of a device-resident network are filled on first use
    const astrosim::Network& host() {
        std::lock_guard<std::mutex> g(mu);
        astrosim::materialize_host_csrs(net);
        return net;
    }
};

struct as_sim {
    std::unique_ptr<astrosim::Engine> eng;
    hipStream_t stream = nullptr;
};

struct as_batch {
    std::unique_ptr<astrosim::Batch> b;
    hipStream_t stream = nullptr;
};

namespace {

template <class F>
as_status guarded(F&& f) {
    try {
        f();
        return AS_OK;
    } catch (const astrosim::Error& e) {
        astrosim::set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        astrosim::set_last_error("host allocation failed");
        return AS_ERR_NOMEM;
    } catch (const std::invalid_argument& e) {
        astrosim::set_last_error(e.what());
        return AS_ERR_INVALID;
    } catch (const std::exception& e) {
        astrosim::set_last_error(e.what());
        return AS_ERR_INVALID;
    }
}

void require(bool ok, const char* msg) {
    if (!ok) throw astrosim::Error(AS_ERR_INVALID, msg);
}

void require_device() {
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        (void)hipGetLastError();
        throw astrosim::Error(AS_ERR_NODEVICE, "No GPU device visible.");
    }
}

std::unique_ptr<astrosim::Engine> make_engine(const astrosim::Network& net, const astrosim::RunParams& rp,
                                              hipStream_t s) {
    if (rp.precision == AS_FP64 && rp.method == AS_METHOD_PLAIN) return astrosim::make_plain_engine(net, rp, s);
    if (rp.method == AS_METHOD_BLOCKED) return astrosim::make_blocked_engine(net, rp, s);
    throw astrosim::Error(AS_ERR_INVALID, "this precision/method combination is not available");
}

}  // namespace

extern "C" {

void as_config_default(as_config* cfg) {
    if (!cfg) return;
    cfg->dt = 0.1;
    cfg->substeps = 1;
    cfg->block = 128;
    cfg->w_n2a = std::nan("");
    cfg->w_a2n = std::nan("");
    cfg->stp_enabled = 1;
    cfg->precision = AS_FP64;
    cfg->method = AS_METHOD_PLAIN;
}

as_status as_network_build(uint32_t n_total, uint64_t seed, as_network** out) {
    return guarded([&] {
        require(out != nullptr, "out is NULL");
        *out = nullptr;
        std::unique_ptr<as_network> n(new as_network);
        astrosim::build_network(n_total, seed, n->net);
        *out = n.release();
    });
}

as_status as_network_build_on(uint32_t n_total, uint64_t seed, int32_t where, as_network** out) {
    return guarded([&] {
        require(out != nullptr, "out is NULL");
        require(where >= AS_BUILD_HOST && where <= AS_BUILD_GPU_ONLY, "where must be 0..2");
        *out = nullptr;
        std::unique_ptr<as_network> n(new as_network);
        std::string why;
        const bool try_gpu = where == AS_BUILD_GPU_ONLY || (where == AS_BUILD_GPU && n_total >= astrosim::kGpuBuildAutoMin);
        if (try_gpu && astrosim::build_network_gpu(n_total, seed, n->net, &why, nullptr, /*keep_on_device=*/true)) {
            n->built_on_gpu = true;
        } else {
            if (where == AS_BUILD_GPU_ONLY)
                throw astrosim::Error(AS_ERR_INVALID, "GPU network build declined: " + why);
            astrosim::build_network(n_total, seed, n->net);
        }
        *out = n.release();
    });
}

as_status as_network_info_get(const as_network* net, as_network_info* info) {
    return guarded([&] {
        require(net && info, "NULL argument");
        const auto& c = net->net.cfg;
        info->n_total = c.n_total;
        info->n_exc = c.n_exc;
        info->n_inh = c.n_inh;
        info->n_astro = c.n_astro;
        info->n_neurons = c.n_neurons();
        info->p_primary = c.p_primary;
        info->n_exc_edges = net->net.edges(0);
        info->n_inh_edges = net->net.edges(1);
        info->n_e2a_edges = net->net.edges(3);
        info->n_sic_edges = net->net.edges(2);
        info->build_seconds = net->net.build_seconds;
        info->built_on_gpu = net->built_on_gpu ? 1 : 0;
    });
}

as_status as_network_csr(const as_network* net, int which, uint32_t* row_start, uint32_t* src) {
    return guarded([&] {
        require(net != nullptr, "net is NULL");
        require(which >= 0 && which < 4, "which must be 0..3");
        const astrosim::Network& hn = const_cast<as_network*>(net)->host();
        const astrosim::HostCSR* c[4] = {&hn.inc_exc, &hn.inc_inh, &hn.inc_sic, &hn.inc_e2a};
        const auto& csr = *c[which];
        if (row_start) std::memcpy(row_start, csr.row_start.data(), csr.row_start.size() * sizeof(uint32_t));
        if (src && !csr.src.empty()) std::memcpy(src, csr.src.data(), csr.src.size() * sizeof(uint32_t));
    });
}

as_status as_network_params(const as_network* net, int which, double* out) {
    return guarded([&] {
        require(net && out, "NULL argument");
        require(which >= 0 && which < 6, "which must be 0..5");
        const std::vector<double>* v[6] = {&net->net.Ca_tot, &net->net.IP3_0, &net->net.tau_IP3,
                                           &net->net.delta_IP3, &net->net.V_reset, &net->net.b};
        std::memcpy(out, v[which]->data(), v[which]->size() * sizeof(double));
    });
}

as_status as_network_dump_edges(const as_network* net, const char* path) {
    return guarded([&] {
        require(net && path, "NULL argument");
        if (!astrosim::dump_edges_csv(const_cast<as_network*>(net)->host(), path))
            throw astrosim::Error(AS_ERR_IO, std::string("cannot open ") + path + " for --dump-edges");
    });
}

void as_network_free(as_network* net) { delete net; }

as_status as_sim_footprint(const as_network* net, const as_config* cfg, uint64_t* bytes) {
    return guarded([&] {
        require(net && cfg && bytes, "NULL argument");
        const astrosim::RunParams rp = astrosim::resolve_params(*cfg);
        *bytes = rp.method == AS_METHOD_PLAIN ? astrosim::plain_footprint(net->net, rp)
                                              : astrosim::blocked_footprint(net->net, rp);
    });
}

as_status as_sim_create(const as_network* net, const as_config* cfg, as_sim** out) {
    return guarded([&] {
        require(net && cfg && out, "NULL argument");
        *out = nullptr;
        const astrosim::RunParams rp = astrosim::resolve_params(*cfg);
        require_device();
        std::unique_ptr<as_sim> s(new as_sim);
        AS_HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        try {
            s->eng = make_engine(net->net, rp, s->stream);
        } catch (...) {
            (void)hipStreamDestroy(s->stream);
            s->stream = nullptr;
            throw;
        }
        *out = s.release();
    });
}

as_status as_sim_step(as_sim* sim, int64_t k) {
    return guarded([&] {
        require(sim != nullptr, "sim is NULL");
        require(k >= 0, "k must be >= 0");
        sim->eng->enqueue_steps(k);
    });
}

as_status as_sim_sync(as_sim* sim) {
    return guarded([&] {
        require(sim != nullptr, "sim is NULL");
        AS_HIP_CHECK(hipStreamSynchronize(sim->stream));
    });
}

int64_t as_sim_steps(const as_sim* sim) { return sim ? sim->eng->steps() : -1; }

as_status as_sim_spike_count(as_sim* sim, uint64_t* out) {
    return guarded([&] {
        require(sim && out, "NULL argument");
        *out = sim->eng->spike_count();
    });
}

as_status as_sim_reset_ever_active(as_sim* sim) {
    return guarded([&] {
        require(sim != nullptr, "sim is NULL");
        sim->eng->reset_ever_active();
    });
}

as_status as_sim_read_state(as_sim* sim, int field, double* out) {
    return guarded([&] {
        require(sim && out, "NULL argument");
        sim->eng->read_state(field, out);
    });
}

as_status as_sim_read_refractory(as_sim* sim, int32_t* out) {
    return guarded([&] {
        require(sim && out, "NULL argument");
        sim->eng->read_refractory(out);
    });
}

as_status as_sim_read_ever_active(as_sim* sim, uint8_t* out) {
    return guarded([&] {
        require(sim && out, "NULL argument");
        sim->eng->read_ever_active(out);
    });
}

as_status as_sim_read_spike_counts(as_sim* sim, uint32_t* out) {
    return guarded([&] {
        require(sim && out, "NULL argument");
        sim->eng->read_spike_counts(out);
    });
}

as_status as_sim_read_spikes(as_sim* sim, int64_t first_step, int32_t n_steps, int64_t* steps, uint32_t* neurons,
                             uint64_t cap, uint64_t* count) {
    return guarded([&] {
        require(sim && count, "NULL argument");
        astrosim::read_spikes(*sim->eng, first_step, n_steps, steps, neurons, cap, count, false);
    });
}

as_status as_sim_read_spikes_noalloc(as_sim* sim, int64_t first_step, int32_t n_steps, int64_t* steps,
                                     uint32_t* neurons, uint64_t cap, uint64_t* count) {
    return guarded([&] {
        require(sim && count, "NULL argument");
        astrosim::read_spikes(*sim->eng, first_step, n_steps, steps, neurons, cap, count, true);
    });
}

as_status as_sim_read_raster(as_sim* sim, int64_t first_step, int32_t n_steps, uint8_t* out) {
    return guarded([&] {
        require(sim && out, "NULL argument");
        sim->eng->read_raster(first_step, n_steps, out);
    });
}

uint64_t as_sim_bytes(const as_sim* sim) { return sim ? sim->eng->device_bytes() : 0; }

as_status as_sim_set_timing(as_sim* sim, int enable) {
    return guarded([&] {
        require(sim != nullptr, "sim is NULL");
        sim->eng->set_timing(enable != 0);
    });
}

as_status as_sim_kernel_time(as_sim* sim, double* total_ms, int64_t* launches) {
    return guarded([&] {
        require(sim && total_ms && launches, "NULL argument");
        sim->eng->kernel_time(total_ms, launches);
    });
}

void as_sim_destroy(as_sim* sim) {
    if (!sim) return;
    sim->eng.reset();
    if (sim->stream) (void)hipStreamDestroy(sim->stream);
    delete sim;
}

static astrosim::BatchCapacity to_cap(const as_batch_capacity& c) {
    astrosim::BatchCapacity k;
    k.neurons = c.neurons;
    k.astros = c.astros;
    for (int i = 0; i < 4; ++i) k.edges[i] = c.edges[i];
    k.members = c.members;
    return k;
}

as_status as_batch_capacity_for(const as_network* const* nets, int32_t n, as_batch_capacity* cap) {
    return guarded([&] {
        require(cap && (n == 0 || nets), "NULL argument");
        require(n >= 0, "n must be >= 0");
        std::vector<const astrosim::Network*> v;
        for (int i = 0; i < n; ++i) {
            require(nets[i] != nullptr, "NULL network");
            v.push_back(&nets[i]->net);
        }
        const astrosim::BatchCapacity k = astrosim::capacity_for(v.data(), n);
        cap->neurons = k.neurons;
        cap->astros = k.astros;
        for (int i = 0; i < 4; ++i) cap->edges[i] = k.edges[i];
        cap->members = k.members;
    });
}

as_status as_batch_footprint(const as_batch_capacity* cap, const as_config* cfg, uint64_t* bytes) {
    return guarded([&] {
        require(cap && cfg && bytes, "NULL argument");
        astrosim::RunParams rp = astrosim::resolve_params(*cfg);
        require(rp.method == AS_METHOD_BLOCKED, "batches run the blocked method");
        *bytes = astrosim::batch_footprint(to_cap(*cap), rp);
    });
}

as_status as_batch_create(const as_batch_capacity* cap, const as_config* cfg, as_batch** out) {
    return guarded([&] {
        require(cap && cfg && out, "NULL argument");
        *out = nullptr;
        const astrosim::RunParams rp = astrosim::resolve_params(*cfg);
        require(rp.method == AS_METHOD_BLOCKED, "batches run the blocked method");
        require(cap->members >= 1, "capacity needs at least one member slot");
        require_device();
        std::unique_ptr<as_batch> b(new as_batch);
        AS_HIP_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        try {
            b->b = astrosim::make_batch(rp, to_cap(*cap), b->stream);
        } catch (...) {
            (void)hipStreamDestroy(b->stream);
            b->stream = nullptr;
            throw;
        }
        *out = b.release();
    });
}

as_status as_batch_fits(as_batch* b, const as_network* net, int32_t* fits) {
    return guarded([&] {
        require(b && net && fits, "NULL argument");
        *fits = b->b->fits(net->net) ? 1 : 0;
    });
}

as_status as_batch_add(as_batch* b, const as_network* net, int64_t n_steps, int32_t* member) {
    return guarded([&] {
        require(b && net && member, "NULL argument");
        *member = b->b->add(net->net, n_steps);
    });
}

as_status as_batch_add_many(as_batch* b, const as_network* const* nets, int32_t n, const int64_t* n_steps,
                            int32_t* members) {
    return guarded([&] {
        require(b && (n == 0 || (nets && n_steps && members)), "NULL argument");
        require(n >= 0, "negative count");
        std::vector<const astrosim::Network*> v(static_cast<std::size_t>(n));
        for (int32_t i = 0; i < n; ++i) {
            require(nets[i] != nullptr, "NULL network");
            v[std::size_t(i)] = &nets[i]->net;
        }
        std::vector<int> m(static_cast<std::size_t>(n), -1);
        try {
            b->b->add_many(v.data(), n, n_steps, m.data());
        } catch (...) {
            for (int32_t i = 0; i < n; ++i) members[i] = m[std::size_t(i)];
            throw;
        }
        for (int32_t i = 0; i < n; ++i) members[i] = m[std::size_t(i)];
    });
}

as_status as_batch_remove(as_batch* b, int32_t member) {
    return guarded([&] {
        require(b != nullptr, "batch is NULL");
        b->b->remove(member);
    });
}

as_status as_batch_step(as_batch* b, int64_t k) {
    return guarded([&] {
        require(b != nullptr, "batch is NULL");
        require(k >= 0, "k must be >= 0");
        b->b->enqueue_steps(k);
    });
}

as_status as_batch_run_graph(as_batch* b, int64_t blocks) {
    return guarded([&] {
        require(b != nullptr, "batch is NULL");
        require(blocks >= 0, "blocks must be >= 0");
        b->b->run_blocks_graph(blocks);
    });
}

as_status as_batch_autotune(as_batch* b, const int32_t* candidates, int32_t n, int32_t blocks_each,
                            int32_t* chosen) {
    return guarded([&] {
        require(b && candidates && chosen, "NULL argument");
        require(n > 0 && blocks_each > 0, "need candidates and blocks_each > 0");
        std::vector<int> c(candidates, candidates + n);
        *chosen = b->b->autotune(c.data(), n, blocks_each);
    });
}

as_status as_batch_sync(as_batch* b) {
    return guarded([&] {
        require(b != nullptr, "batch is NULL");
        AS_HIP_CHECK(hipStreamSynchronize(b->stream));
    });
}

as_status as_batch_get_info(as_batch* b, as_batch_info* info) {
    return guarded([&] {
        require(b && info, "NULL argument");
        info->global_step = b->b->global_step();
        info->block_steps = b->b->block_len();
        info->active_members = b->b->active_members();
        info->block_threads = b->b->block_threads();
        info->device_bytes = b->b->device_bytes();
        info->free_neurons = b->b->free_neurons();
        info->bucket_overflows = b->b->member_overflows();
        info->chunk_steps = b->b->chunk_steps();
    });
}

as_status as_batch_member_steps(as_batch* b, int32_t member, int64_t* done, int64_t* target) {
    return guarded([&] {
        require(b && done && target, "NULL argument");
        *done = b->b->member_steps(member);
        *target = b->b->member_target(member);
    });
}

as_status as_batch_spike_count(as_batch* b, int32_t member, uint64_t* out) {
    return guarded([&] {
        require(b && out, "NULL argument");
        *out = b->b->spike_count(member);
    });
}

as_status as_batch_reset_ever_active(as_batch* b, int32_t member) {
    return guarded([&] {
        require(b != nullptr, "batch is NULL");
        b->b->reset_ever_active(member);
    });
}

as_status as_batch_read_state(as_batch* b, int32_t member, int32_t field, double* out) {
    return guarded([&] {
        require(b && out, "NULL argument");
        b->b->read_state(member, field, out);
    });
}

as_status as_batch_read_ever_active(as_batch* b, int32_t member, uint8_t* out) {
    return guarded([&] {
        require(b && out, "NULL argument");
        b->b->read_ever_active(member, out);
    });
}

as_status as_batch_read_spike_counts(as_batch* b, int32_t member, uint32_t* out) {
    return guarded([&] {
        require(b && out, "NULL argument");
        b->b->read_spike_counts(member, out);
    });
}

as_status as_batch_read_raster(as_batch* b, int32_t member, int64_t first_step, int32_t n_steps, uint8_t* out) {
    return guarded([&] {
        require(b && out, "NULL argument");
        b->b->read_raster(member, first_step, n_steps, out);
    });
}

as_status as_batch_set_timing(as_batch* b, int enable) {
    return guarded([&] {
        require(b != nullptr, "batch is NULL");
        b->b->set_timing(enable != 0);
    });
}

as_status as_batch_kernel_time(as_batch* b, double* total_ms, int64_t* launches) {
    return guarded([&] {
        require(b && total_ms && launches, "NULL argument");
        b->b->kernel_time(total_ms, launches);
    });
}

void as_batch_destroy(as_batch* b) {
    if (!b) return;
    b->b.reset();
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

as_status as_device_info(char* name, int name_len, char* arch, int arch_len) {
    return guarded([&] {
        require_device();
        hipDeviceProp_t prop{};
        AS_HIP_CHECK(hipGetDeviceProperties(&prop, 0));
        if (name && name_len > 0) {
            std::strncpy(name, prop.name, name_len - 1);
            name[name_len - 1] = '\0';
        }
        if (arch && arch_len > 0) {
            std::strncpy(arch, prop.gcnArchName, arch_len - 1);
            arch[arch_len - 1] = '\0';
        }
    });
}

as_status as_device_memory(uint64_t* free_bytes, uint64_t* total_bytes) {
    return guarded([&] {
        require_device();
        size_t f = 0, t = 0;
        AS_HIP_CHECK(hipMemGetInfo(&f, &t));
        if (free_bytes) *free_bytes = f;
        if (total_bytes) *total_bytes = t;
    });
}

const char* as_last_error(void) { return astrosim::last_error(); }

const char* as_version(void) { return "astrosim-mi355x 0.1.0 (gfx950)"; }

}  // extern "C"
