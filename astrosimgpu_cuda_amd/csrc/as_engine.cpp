// as_engine.cpp -- parameter resolution and the engine-independent parts.
#include "as_engine.h"

#include <cmath>
#include <cstdio>

namespace astrosim {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

void throw_hip(hipError_t e, const char* file, int line, const char* expr) {
    char buf[512];
    std::snprintf(buf, sizeof buf, "HIP error %s:%d: %s (%s)", file, line, hipGetErrorString(e), expr);
    (void)hipGetLastError();
    throw Error(AS_ERR_HIP, buf);
}

RunParams resolve_params(const as_config& c) {
    if (!(c.dt > 0.0) || !std::isfinite(c.dt)) throw Error(AS_ERR_INVALID, "dt must be a positive number");
    if (c.substeps < 1) throw Error(AS_ERR_INVALID, "substeps must be >= 1");
    if (c.block < 64 || c.block > 1024 || c.block % 64 != 0)
        throw Error(AS_ERR_INVALID, "block must be a multiple of 64 in [64, 1024] (wave64)");
    if (c.precision != AS_FP64 && c.precision != AS_FP32C) throw Error(AS_ERR_INVALID, "unknown precision");
    if (c.method != AS_METHOD_PLAIN && c.method != AS_METHOD_BLOCKED) throw Error(AS_ERR_INVALID, "unknown method");
    Biology bio;
    // Overrides and noise_dt resolution (src/astrosim_fused.cu:847-851).
    if (bio.noise_dt <= 0.0) bio.noise_dt = 10.0 * c.dt;
    if (!std::isnan(c.w_n2a)) bio.w_n2a = c.w_n2a;
    if (!std::isnan(c.w_a2n)) bio.w_a2n = c.w_a2n;
    if (!c.stp_enabled) bio.stp_enabled = false;

    RunParams P;
    P.dt = c.dt;
    P.substeps = c.substeps;
    P.h_step = c.dt / static_cast<double>(c.substeps);  // src:1024
    P.h_sixth = P.h_step / 6.0;
    P.block = c.block;
    P.precision = c.precision;
    P.method = c.method;
    P.w_e = bio.w_e;
    P.w_i = bio.w_i;
    P.w_n2a = bio.w_n2a;
    P.w_a2n = bio.w_a2n;
    P.stp_enabled = bio.stp_enabled;
    P.stp_U = bio.stp_U;
    P.stp_tau_rec = bio.stp_tau_rec;
    P.stp_tau_fac = bio.stp_tau_fac;
    // Delays and ring depth (src:954-957); noise refresh period (src:1043).
    P.d_e = delay_to_steps(bio.d_e, c.dt);
    P.d_i = delay_to_steps(bio.d_i, c.dt);
    P.d_a2n = delay_to_steps(bio.d_a2n, c.dt);
    P.ring_slots = std::max(std::max(P.d_e, P.d_i), P.d_a2n) + 1;
    const long long spn = std::llround(bio.noise_dt / c.dt);
    P.steps_per_noise = spn < 1 ? 1 : static_cast<int>(spn);
    P.sic_interval = bio.sic_interval;
    // Poisson drive (src:393-398, src:472-479): lambda = rate*dt*1e-3.
    P.poiss_astro = bio.astro_poiss_rate > 0.0;
    P.lam_astro = bio.astro_poiss_rate * c.dt * 1e-3;
    P.lim_astro = as_exp(-P.lam_astro);
    P.w_poiss_astro = bio.astro_poiss_weight;
    P.poiss_exc = bio.exc_poiss_rate > 0.0;
    P.lam_exc = bio.exc_poiss_rate * c.dt * 1e-3;
    P.lim_exc = as_exp(-P.lam_exc);
    P.w_poiss_exc = bio.exc_poiss_weight;
    P.poiss_inh = bio.inh_poiss_rate > 0.0;
    P.lam_inh = bio.inh_poiss_rate * c.dt * 1e-3;
    P.lim_inh = as_exp(-P.lam_inh);
    P.w_poiss_inh = bio.inh_poiss_weight;
    // Independent Gaussian noise (the configuration's independent_noise=true).
    P.noise_astro = bio.astro_noise_std;
    P.noise_exc = bio.exc_noise_std;
    P.noise_inh = bio.inh_noise_std;
    P.ac = make_astro_const(bio);
    P.ne = make_neuron_const(bio, true, P.h_step);
    P.ni = make_neuron_const(bio, false, P.h_step);
    // Alpha-cascade propagators, per population (src:515-516).
    P.ex_decay_e = as_exp(-P.h_step / P.ne.tau_syn_ex);
    P.in_decay_e = as_exp(-P.h_step / P.ne.tau_syn_in);
    P.ex_decay_i = as_exp(-P.h_step / P.ni.tau_syn_ex);
    P.in_decay_i = as_exp(-P.h_step / P.ni.tau_syn_in);
    P.SIC_th = bio.SIC_th;
    return P;
}

void Engine::set_timing(bool on) { timing_ = on; }

void Engine::timing_begin() {
    if (!timing_) return;
    AS_HIP_CHECK(hipEventCreate(&pending_start_));
    AS_HIP_CHECK(hipEventRecord(pending_start_, stream_));
}

void Engine::timing_end(std::int64_t launches) {
    if (!timing_ || !pending_start_) return;
    hipEvent_t stop;
    AS_HIP_CHECK(hipEventCreate(&stop));
    AS_HIP_CHECK(hipEventRecord(stop, stream_));
    events_.emplace_back(pending_start_, stop);
    pending_start_ = nullptr;
    timed_launches_ += launches;
}

void Engine::kernel_time(double* total_ms, std::int64_t* launches) {
    AS_HIP_CHECK(hipStreamSynchronize(stream_));
    double t = 0.0;
    for (auto& e : events_) {
        float ms = 0.0f;
        AS_HIP_CHECK(hipEventElapsedTime(&ms, e.first, e.second));
        t += ms;
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    events_.clear();
    *total_ms = t;
    *launches = timed_launches_;
    timed_launches_ = 0;
}

}  // namespace astrosim
