// as_plain.hip -- the plain FP64 method: one fused launch per step.
//
// Data flow follows the reference (src/astrosim_fused.cu:374-587): dense
// delay rings [R][cells] of doubles, one lane per cell, each cell gathering
// its in-edges at the delayed slot and writing only its own state and its
// own ring slot. The astrocyte and neuron phases of a step read only slots
// written in earlier steps (every delay is >= 1 step and < R), so both run
// in ONE launch: workgroups [0, astro_blocks) take astrocytes, the rest take
// neurons. This is the FP64 baseline the blocked method is measured against.
#include "as_engine.h"
#include "as_model.h"

namespace astrosim {
namespace {

struct PlainDev {
    const index_t *exc_rs, *exc_src, *inh_rs, *inh_src, *sic_rs, *sic_src, *e2a_rs, *e2a_src;
    double *Ca, *IP3, *h;
    const double *Ca_tot, *IP3_0, *tau_IP3, *delta_IP3;
    std::uint8_t* ever;
    double *V, *w, *g_ex, *dg_ex, *g_in, *dg_in, *I_sic;
    std::int32_t* refr;
    const double *V_reset, *b;
    double *stp_x, *stp_u, *stp_t;
    double *ring_n, *ring_s;  // [R][n_neurons], [R][n_astro]
    unsigned long long* counters;  // kSpikeCounters partial spike totals
    index_t n_astro, n_neurons, n_exc;
    std::uint32_t astro_blocks;
};

struct PlainStep {
    std::int64_t step;
    std::uint64_t noise_idx;
    int sic_refresh;
    std::uint32_t slot_ex, slot_in, slot_sic, slot_w;
};

constexpr int kSpikeCounters = 64;

template <int MAXT>
__global__ __launch_bounds__(MAXT) void plain_step_kernel(PlainDev D, RunParams P,
                                                          std::uint64_t seed, PlainStep S) {
    if (blockIdx.x < D.astro_blocks) {
        // ---- astrocyte phase (src/astrosim_fused.cu:374-427) ----
        const index_t a = blockIdx.x * blockDim.x + threadIdx.x;
        if (a >= D.n_astro) return;
        const double* row = D.ring_n + std::size_t(S.slot_ex) * D.n_neurons;
        double gathered = 0.0;
        for (index_t k = D.e2a_rs[a]; k < D.e2a_rs[a + 1]; ++k) gathered += row[D.e2a_src[k]];
        double ip3_input = P.w_n2a * gathered;
        const std::uint64_t stream = std::uint64_t(S.step) * kStepStride + a;
        if (P.poiss_astro) {
            const int ev = rng_poisson(seed ^ kKeyAstroPoisson, stream, P.lam_astro, P.lim_astro);
            if (ev > 0) ip3_input += P.w_poiss_astro * static_cast<double>(ev);
        }
        const double noise =
            P.noise_astro > 0.0 ? P.noise_astro * rng_normal(seed ^ kKeyAstroNoise, stream) : 0.0;
        double ca = D.Ca[a], ip3 = D.IP3[a], hh = D.h[a];
        astro_advance(P.ac, D.Ca_tot[a], D.IP3_0[a], make_fixed_divisor(D.tau_IP3[a]), D.delta_IP3[a],
                      ip3_input, noise, P.h_step, P.h_sixth, P.substeps, ca, ip3, hh);
        D.Ca[a] = ca;
        D.IP3[a] = ip3;
        D.h[a] = hh;
        const double factor = sic_factor(P.ac, ca);
        if (factor > 0.0) D.ever[a] = 1;
        D.ring_s[std::size_t(S.slot_w) * D.n_astro + a] = factor;
        return;
    }

    // ---- neuron phase (src/astrosim_fused.cu:433-587) ----
    const index_t j = (blockIdx.x - D.astro_blocks) * blockDim.x + threadIdx.x;
    bool spiked = false;
    if (j < D.n_neurons) {
        const bool exc = j < D.n_exc;
        const NeuronConst& p = exc ? P.ne : P.ni;
        const double* row_ex = D.ring_n + std::size_t(S.slot_ex) * D.n_neurons;
        double exc_sum = 0.0;
        for (index_t k = D.exc_rs[j]; k < D.exc_rs[j + 1]; ++k) exc_sum += row_ex[D.exc_src[k]];
        // Inhibitory source ids are local indices and index the neuron ring
        // as they are (src/astrosim_fused.cu:459-462).
        const double* row_in = D.ring_n + std::size_t(S.slot_in) * D.n_neurons;
        double inh_sum = 0.0;
        for (index_t k = D.inh_rs[j]; k < D.inh_rs[j + 1]; ++k) inh_sum += row_in[D.inh_src[k]];
        double exc_input = P.w_e * exc_sum;
        const double inh_input = (-P.w_i) * inh_sum;
        if (exc ? P.poiss_exc : P.poiss_inh) {
            const int ev = rng_poisson(seed ^ kKeyNeuronPoisson, std::uint64_t(S.step) * kStepStride + j,
                                       exc ? P.lam_exc : P.lam_inh, exc ? P.lim_exc : P.lim_inh);
            if (ev > 0) exc_input += (exc ? P.w_poiss_exc : P.w_poiss_inh) * static_cast<double>(ev);
        }
        double I_sic;
        if (S.sic_refresh) {
            const double* row_sic = D.ring_s + std::size_t(S.slot_sic) * D.n_astro;
            double sic_sum = 0.0;
            for (index_t k = D.sic_rs[j]; k < D.sic_rs[j + 1]; ++k) sic_sum += row_sic[D.sic_src[k]];
            I_sic = P.w_a2n * sic_sum;
            D.I_sic[j] = I_sic;
        } else {
            I_sic = D.I_sic[j];
        }
        double I_ext = I_sic;
        const double nstd = exc ? P.noise_exc : P.noise_inh;
        if (nstd > 0.0) {
            I_ext += nstd * rng_normal(seed ^ kKeyNeuronNoise, S.noise_idx * kNoiseStride + std::uint64_t(j));
        }
        NeuronRegs r;
        r.V = D.V[j];
        r.w = D.w[j];
        r.g_ex = D.g_ex[j];
        r.dg_ex = D.dg_ex[j] + exc_input * p.psc_init_ex;
        r.g_in = D.g_in[j];
        r.dg_in = D.dg_in[j] + inh_input * p.psc_init_in;
        r.refractory = D.refr[j];
        spiked = neuron_substeps(p, r, I_ext, D.V_reset[j], D.b[j], P.h_step, P.h_sixth, P.substeps,
                                 exc ? P.ex_decay_e : P.ex_decay_i, exc ? P.in_decay_e : P.in_decay_i);
        D.V[j] = r.V;
        D.w[j] = r.w;
        D.g_ex[j] = r.g_ex;
        D.dg_ex[j] = r.dg_ex;
        D.g_in[j] = r.g_in;
        D.dg_in[j] = r.dg_in;
        D.refr[j] = r.refractory;
        double out = 0.0;
        if (spiked) {
            if (P.stp_enabled) {
                double x = D.stp_x[j], u = D.stp_u[j], tl = D.stp_t[j];
                out = stp_on_spike(x, u, tl, static_cast<double>(S.step) * P.dt, P.stp_U,
                                   P.stp_tau_rec, P.stp_tau_fac);
                D.stp_x[j] = x;
                D.stp_u[j] = u;
                D.stp_t[j] = tl;
            } else {
                out = 1.0;
            }
        }
        D.ring_n[std::size_t(S.slot_w) * D.n_neurons + j] = out;
    }
    // Diagnostic spike count: one atomic per wave that saw a spike, spread
    // over kSpikeCounters addresses (reference: one atomicAdd per spike, src:584).
    const unsigned long long m = __ballot(spiked);
    if ((threadIdx.x & 63u) == 0 && m != 0ull)
        atomicAdd(&D.counters[blockIdx.x & (kSpikeCounters - 1)], (unsigned long long)__popcll(m));
}

__global__ void fill_kernel(double* p, double v, std::size_t n) {
    const std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

struct PlainLayout {
    std::size_t csr_rs[4], csr_src[4];
    std::size_t astro[7], ever;
    std::size_t neuron[12], refr;
    std::size_t ring_n, ring_s, counters;
    std::size_t total;
};

PlainLayout plain_layout(const Network& net, const RunParams& rp) {
    Layout L;
    PlainLayout o{};
    const HostCSR* c[4] = {&net.inc_exc, &net.inc_inh, &net.inc_sic, &net.inc_e2a};
    for (int i = 0; i < 4; ++i) {
        const std::size_t rows = net.on_device() ? std::size_t(net.dev[i].n_targets) + 1 : c[i]->row_start.size();
        o.csr_rs[i] = L.add(rows * sizeof(index_t));
        o.csr_src[i] = L.add(std::max<std::size_t>(1, net.edges(i)) * sizeof(index_t));
    }
    const std::size_t na = net.cfg.n_astro, nn = net.cfg.n_neurons();
    for (int i = 0; i < 7; ++i) o.astro[i] = L.add(na * sizeof(double));
    o.ever = L.add(na);
    for (int i = 0; i < 12; ++i) o.neuron[i] = L.add(nn * sizeof(double));
    o.refr = L.add(nn * sizeof(std::int32_t));
    o.ring_n = L.add(std::size_t(rp.ring_slots) * nn * sizeof(double));
    o.ring_s = L.add(std::size_t(rp.ring_slots) * na * sizeof(double));
    o.counters = L.add(kSpikeCounters * sizeof(unsigned long long));
    o.total = L.total;
    return o;
}

class PlainEngine final : public Engine {
public:
    PlainEngine(const Network& net, const RunParams& rp, hipStream_t s) : P_(rp), seed_(net.seed) {
        stream_ = s;
        n_neurons_ = net.cfg.n_neurons();
        n_astro_ = net.cfg.n_astro;
        n_exc_ = net.cfg.n_exc;
        lay_ = plain_layout(net, rp);
        if (hipMalloc(&base_, lay_.total) != hipSuccess) {
            (void)hipGetLastError();
            base_ = nullptr;
            throw Error(AS_ERR_NOMEM, "device allocation of " + std::to_string(lay_.total) + " bytes failed");
        }
        char* b = static_cast<char*>(base_);
        auto dptr = [&](std::size_t off) { return static_cast<void*>(b + off); };
        const HostCSR* c[4] = {&net.inc_exc, &net.inc_inh, &net.inc_sic, &net.inc_e2a};
        const bool dev = net.on_device();  // a network generated on this GPU: device to device
        for (int i = 0; i < 4; ++i) {
            const std::size_t rows = dev ? std::size_t(net.dev[i].n_targets) + 1 : c[i]->row_start.size();
            AS_HIP_CHECK(hipMemcpy(dptr(lay_.csr_rs[i]), dev ? net.dev[i].row_start : c[i]->row_start.data(),
                                   rows * sizeof(index_t), dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
            if (net.edges(i))
                AS_HIP_CHECK(hipMemcpy(dptr(lay_.csr_src[i]), dev ? net.dev[i].src : c[i]->src.data(),
                                       net.edges(i) * sizeof(index_t),
                                       dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice));
        }
        D_.exc_rs = (const index_t*)dptr(lay_.csr_rs[0]);
        D_.exc_src = (const index_t*)dptr(lay_.csr_src[0]);
        D_.inh_rs = (const index_t*)dptr(lay_.csr_rs[1]);
        D_.inh_src = (const index_t*)dptr(lay_.csr_src[1]);
        D_.sic_rs = (const index_t*)dptr(lay_.csr_rs[2]);
        D_.sic_src = (const index_t*)dptr(lay_.csr_src[2]);
        D_.e2a_rs = (const index_t*)dptr(lay_.csr_rs[3]);
        D_.e2a_src = (const index_t*)dptr(lay_.csr_src[3]);
        double* A[7];
        for (int i = 0; i < 7; ++i) A[i] = (double*)dptr(lay_.astro[i]);
        D_.Ca = A[0]; D_.IP3 = A[1]; D_.h = A[2];
        D_.Ca_tot = A[3]; D_.IP3_0 = A[4]; D_.tau_IP3 = A[5]; D_.delta_IP3 = A[6];
        D_.ever = (std::uint8_t*)dptr(lay_.ever);
        double* N[12];
        for (int i = 0; i < 12; ++i) N[i] = (double*)dptr(lay_.neuron[i]);
        D_.V = N[0]; D_.w = N[1]; D_.g_ex = N[2]; D_.dg_ex = N[3]; D_.g_in = N[4]; D_.dg_in = N[5];
        D_.I_sic = N[6]; D_.stp_x = N[7]; D_.stp_u = N[8]; D_.stp_t = N[9];
        D_.V_reset = N[10]; D_.b = N[11];
        D_.refr = (std::int32_t*)dptr(lay_.refr);
        D_.ring_n = (double*)dptr(lay_.ring_n);
        D_.ring_s = (double*)dptr(lay_.ring_s);
        D_.counters = (unsigned long long*)dptr(lay_.counters);
        D_.n_astro = n_astro_;
        D_.n_neurons = n_neurons_;
        D_.n_exc = n_exc_;
        D_.astro_blocks = (n_astro_ + rp.block - 1) / rp.block;
        neuron_blocks_ = (n_neurons_ + rp.block - 1) / rp.block;

        // Parameters and initial state (src/astrosim_fused.cu:981-1018).
        const Biology bio;
        AS_HIP_CHECK(hipMemcpy(A[3], net.Ca_tot.data(), n_astro_ * sizeof(double), hipMemcpyHostToDevice));
        AS_HIP_CHECK(hipMemcpy(A[4], net.IP3_0.data(), n_astro_ * sizeof(double), hipMemcpyHostToDevice));
        AS_HIP_CHECK(hipMemcpy(A[5], net.tau_IP3.data(), n_astro_ * sizeof(double), hipMemcpyHostToDevice));
        AS_HIP_CHECK(hipMemcpy(A[6], net.delta_IP3.data(), n_astro_ * sizeof(double), hipMemcpyHostToDevice));
        AS_HIP_CHECK(hipMemcpy(N[10], net.V_reset.data(), n_neurons_ * sizeof(double), hipMemcpyHostToDevice));
        AS_HIP_CHECK(hipMemcpy(N[11], net.b.data(), n_neurons_ * sizeof(double), hipMemcpyHostToDevice));
        fill(A[0], bio.Ca_init, n_astro_);
        fill(A[1], bio.IP3_init, n_astro_);
        fill(A[2], bio.h_init, n_astro_);
        {
            std::vector<double> V0(n_neurons_);
            for (index_t i = 0; i < n_neurons_; ++i) V0[i] = i < n_exc_ ? bio.e_V_m : bio.i_V_m;
            AS_HIP_CHECK(hipMemcpy(N[0], V0.data(), n_neurons_ * sizeof(double), hipMemcpyHostToDevice));
        }
        for (int i = 1; i <= 6; ++i) AS_HIP_CHECK(hipMemsetAsync(N[i], 0, n_neurons_ * sizeof(double), stream_));
        fill(N[7], 1.0, n_neurons_);
        fill(N[8], rp.stp_U, n_neurons_);
        fill(N[9], -1e12, n_neurons_);
        AS_HIP_CHECK(hipMemsetAsync(D_.refr, 0, n_neurons_ * sizeof(std::int32_t), stream_));
        AS_HIP_CHECK(hipMemsetAsync(D_.ever, 0, n_astro_, stream_));
        AS_HIP_CHECK(hipMemsetAsync(D_.ring_n, 0, std::size_t(rp.ring_slots) * n_neurons_ * sizeof(double), stream_));
        AS_HIP_CHECK(hipMemsetAsync(D_.ring_s, 0, std::size_t(rp.ring_slots) * n_astro_ * sizeof(double), stream_));
        AS_HIP_CHECK(hipMemsetAsync(D_.counters, 0, kSpikeCounters * sizeof(unsigned long long), stream_));
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    ~PlainEngine() override {
        if (base_) {
            (void)hipStreamSynchronize(stream_);
            (void)hipFree(base_);
        }
    }

    void enqueue_steps(std::int64_t k) override {
        if (k <= 0) return;
        const int R = P_.ring_slots;
        const unsigned grid = D_.astro_blocks + neuron_blocks_;
        for (std::int64_t i = 0; i < k; ++i) {
            const std::int64_t step = step_ + i;
            PlainStep S;
            S.step = step;
            S.noise_idx = static_cast<std::uint64_t>(step / P_.steps_per_noise);
            S.sic_refresh = (step - P_.d_a2n >= 0) && ((step - P_.d_a2n) % P_.sic_interval == 0);
            S.slot_ex = static_cast<std::uint32_t>((step + R - P_.d_e) % R);
            S.slot_in = static_cast<std::uint32_t>((step + R - P_.d_i) % R);
            S.slot_sic = static_cast<std::uint32_t>((step + R - P_.d_a2n) % R);
            S.slot_w = static_cast<std::uint32_t>(step % R);
            timing_begin();
            if (P_.block <= 256)
                hipLaunchKernelGGL(plain_step_kernel<256>, dim3(grid), dim3(P_.block), 0, stream_, D_, P_, seed_, S);
            else
                hipLaunchKernelGGL(plain_step_kernel<1024>, dim3(grid), dim3(P_.block), 0, stream_, D_, P_, seed_, S);
            timing_end(1);
        }
        AS_HIP_CHECK(hipGetLastError());
        step_ += k;
    }

    std::uint64_t spike_count() override {
        unsigned long long h[kSpikeCounters];
        AS_HIP_CHECK(hipMemcpyAsync(h, D_.counters, sizeof h, hipMemcpyDeviceToHost, stream_));
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
        std::uint64_t t = 0;
        for (int i = 0; i < kSpikeCounters; ++i) t += h[i];
        return t;
    }

    void reset_ever_active() override {
        AS_HIP_CHECK(hipMemsetAsync(D_.ever, 0, n_astro_, stream_));
    }

    void read_state(int field, double* out) override {
        const double* src = nullptr;
        std::size_t n = n_neurons_;
        switch (field) {
            case 0: src = D_.Ca; n = n_astro_; break;
            case 1: src = D_.IP3; n = n_astro_; break;
            case 2: src = D_.h; n = n_astro_; break;
            case 3: src = D_.V; break;
            case 4: src = D_.w; break;
            case 5: src = D_.g_ex; break;
            case 6: src = D_.dg_ex; break;
            case 7: src = D_.g_in; break;
            case 8: src = D_.dg_in; break;
            case 9: src = D_.I_sic; break;
            case 10: src = D_.stp_x; break;
            case 11: src = D_.stp_u; break;
            case 12: src = D_.stp_t; break;
            default: throw Error(AS_ERR_INVALID, "unknown state field " + std::to_string(field));
        }
        AS_HIP_CHECK(hipMemcpyAsync(out, src, n * sizeof(double), hipMemcpyDeviceToHost, stream_));
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void read_refractory(std::int32_t* out) override {
        AS_HIP_CHECK(hipMemcpyAsync(out, D_.refr, n_neurons_ * sizeof(std::int32_t), hipMemcpyDeviceToHost, stream_));
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void read_ever_active(std::uint8_t* out) override {
        AS_HIP_CHECK(hipMemcpyAsync(out, D_.ever, n_astro_, hipMemcpyDeviceToHost, stream_));
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void read_raster(std::int64_t first, int n, std::uint8_t* out) override {
        // Slot s % R holds step s for the last R steps.
        const int R = P_.ring_slots;
        if (n < 0 || first < 0 || first + n > step_ || first < step_ - R)
            throw Error(AS_ERR_RANGE, "raster steps not held (the plain method keeps the last R steps)");
        std::vector<double> row(n_neurons_);
        for (int i = 0; i < n; ++i) {
            const std::int64_t s = first + i;
            AS_HIP_CHECK(hipMemcpyAsync(row.data(), D_.ring_n + std::size_t(s % R) * n_neurons_,
                                        n_neurons_ * sizeof(double), hipMemcpyDeviceToHost, stream_));
            AS_HIP_CHECK(hipStreamSynchronize(stream_));
            for (index_t j = 0; j < n_neurons_; ++j) out[std::size_t(i) * n_neurons_ + j] = row[j] != 0.0;
        }
    }

    std::uint64_t device_bytes() const override { return lay_.total; }

private:
    void fill(double* p, double v, std::size_t n) {
        if (n == 0) return;
        hipLaunchKernelGGL(fill_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream_, p, v, n);
        AS_HIP_CHECK(hipGetLastError());
    }

    RunParams P_;
    std::uint64_t seed_;
    PlainDev D_{};
    PlainLayout lay_{};
    void* base_ = nullptr;
    std::uint32_t neuron_blocks_ = 0;
};

}  // namespace

std::uint64_t plain_footprint(const Network& net, const RunParams& rp) { return plain_layout(net, rp).total; }

std::unique_ptr<Engine> make_plain_engine(const Network& net, const RunParams& rp, hipStream_t s) {
    return std::unique_ptr<Engine>(new PlainEngine(net, rp, s));
}

}  // namespace astrosim
