// as_setup.cpp -- host network construction (see as_setup.h).
#include "as_setup.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <thread>

namespace astrosim {

ScalingConfig make_scaling(index_t n_total) {
    // Reference: src/astrosim_fused.cu:606-614.
    ScalingConfig cfg;
    cfg.n_total = n_total;
    cfg.n_exc = std::max<index_t>(1, static_cast<index_t>(std::lround(n_total * ScalingConfig::kEFraction)));
    cfg.n_inh = std::max<index_t>(1, n_total - cfg.n_exc);
    cfg.n_astro =
        std::max<index_t>(1, static_cast<index_t>(std::lround(n_total / ScalingConfig::kAstroRatio)));
    cfg.p_primary = std::min<double>(1.0, ScalingConfig::kSyn / static_cast<double>(cfg.n_exc));
    return cfg;
}

AstroConst make_astro_const(const Biology& bio) {
    // Reference: src/astrosim_fused.cu:935-938.
    AstroConst c{};
    c.Kd_IP3_1 = bio.Kd_IP3_1;
    c.Kd_IP3_2 = bio.Kd_IP3_2;
    c.Kd_act = bio.Kd_act;
    c.Kd_inh = bio.Kd_inh;
    c.Km_SERCA = bio.Km_SERCA;
    c.k_IP3R = bio.k_IP3R;
    c.rate_IP3R = bio.rate_IP3R;
    c.rate_L = bio.rate_L;
    c.rate_SERCA = bio.rate_SERCA;
    c.ratio_ER_cyt = bio.ratio_ER_cyt;
    c.SIC_scale = bio.SIC_scale;
    c.SIC_th = bio.SIC_th;
    return c;
}

NeuronConst make_neuron_const(const Biology& bio, bool exc) {
    // Reference: src/astrosim_fused.cu:940-951. psc_init = e / tau_syn.
    NeuronConst p{};
    p.C_m = exc ? bio.e_C_m : bio.i_C_m;
    p.g_L = exc ? bio.e_g_L : bio.i_g_L;
    p.E_L = exc ? bio.e_E_L : bio.i_E_L;
    p.V_th = exc ? bio.e_V_th : bio.i_V_th;
    p.Delta_T = exc ? bio.e_Delta_T : bio.i_Delta_T;
    p.a = exc ? bio.e_a : bio.i_a;
    p.tau_w = exc ? bio.e_tau_w : bio.i_tau_w;
    p.V_peak = exc ? bio.e_V_peak : bio.i_V_peak;
    p.t_ref = exc ? bio.e_t_ref : bio.i_t_ref;
    p.E_ex = exc ? bio.e_E_ex : bio.i_E_ex;
    p.E_in = exc ? bio.e_E_in : bio.i_E_in;
    p.tau_syn_ex = exc ? bio.e_tau_syn_ex : bio.i_tau_syn_ex;
    p.tau_syn_in = exc ? bio.e_tau_syn_in : bio.i_tau_syn_in;
    p.I_e = 0.0;
    p.psc_init_ex = 2.7182818284590452353602874713527 / p.tau_syn_ex;
    p.psc_init_in = 2.7182818284590452353602874713527 / p.tau_syn_in;
    return p;
}

double HostRng::normal() {
    if (has_spare_) {
        has_spare_ = false;
        return spare_;
    }
    double u1;
    do {
        u1 = uniform();
    } while (u1 <= 1e-300);
    const double u2 = uniform();
    const double r = std::sqrt(-2.0 * std::log(u1));
    const double theta = 6.283185307179586476925286766559 * u2;
    // sin and cos of one angle through the C library's sincos, which is what
    // an optimising host compiler emits for the reference's two calls; the
    // separate sin/cos entry points round differently in rare cases.
    double sn, cs;
    ::sincos(theta, &sn, &cs);
    spare_ = r * sn;
    has_spare_ = true;
    return r * cs;
}

double HostRng::normal_redraw(double mean, double stddev, double lo, double hi, int max_tries) {
    for (int i = 0; i < max_tries; ++i) {
        const double v = mean + stddev * normal();
        if (v >= lo && v <= hi) return v;
    }
    return std::min(std::max(mean, lo), hi);
}

double draw_scaled(double mean, bool enabled, double lower, double upper, double var, HostRng& rng) {
    if (!enabled || var <= 0.0) return mean;
    const double lo = mean < 0.0 ? mean * upper : mean * lower;
    const double hi = mean < 0.0 ? mean * lower : mean * upper;
    return rng.normal_redraw(mean, var * std::fabs(mean), lo, hi);
}

index_t geometric_skip(double p, double log1m_p, index_t remaining, HostRng& rng) {
    if (remaining == 0 || p >= 1.0) return 0;
    if (p <= 0.0) return remaining;
    double u = rng.uniform();
    if (u >= 1.0) u = std::nextafter(1.0, 0.0);
    const double g = std::floor(std::log1p(-u) / log1m_p);
    return (g >= static_cast<double>(remaining)) ? remaining : static_cast<index_t>(g);
}

namespace {

// Edges of one generation pass kept compactly: the targets of source `s` are
// dst[off[s] .. off[s+1]), in generation order.
struct SourceMajor {
    std::vector<index_t> dst;
    std::vector<std::size_t> off;
};

// Stable counting-sort inversion of source-major passes into a per-target CSR.
// The passes are consumed in order, so row t lists its sources exactly in
// generation order. `src_base[k]` is added to source ids of pass k.
void invert(index_t n_targets, const std::vector<const SourceMajor*>& passes,
            const std::vector<index_t>& indeg, HostCSR& out) {
    out.row_start.assign(std::size_t(n_targets) + 1, 0);
    for (index_t t = 0; t < n_targets; ++t) out.row_start[t + 1] = out.row_start[t] + indeg[t];
    out.src.assign(out.row_start[n_targets], 0);
    std::vector<index_t> cursor(out.row_start.begin(), out.row_start.end() - 1);
    for (const SourceMajor* pass : passes) {
        const std::size_t n_src = pass->off.size() - 1;
        for (std::size_t s = 0; s < n_src; ++s) {
            for (std::size_t k = pass->off[s]; k < pass->off[s + 1]; ++k) {
                out.src[cursor[pass->dst[k]]++] = static_cast<index_t>(s);
            }
        }
    }
}

// Pairs variant for the third-factor edges (source ids are explicit).
void invert_pairs(index_t n_targets, const std::vector<index_t>& srcs,
                  const std::vector<index_t>& dsts, HostCSR& out) {
    out.row_start.assign(std::size_t(n_targets) + 1, 0);
    for (index_t d : dsts) out.row_start[std::size_t(d) + 1]++;
    for (index_t t = 0; t < n_targets; ++t) out.row_start[t + 1] += out.row_start[t];
    out.src.assign(srcs.size(), 0);
    std::vector<index_t> cursor(out.row_start.begin(), out.row_start.end() - 1);
    for (std::size_t k = 0; k < srcs.size(); ++k) out.src[cursor[dsts[k]]++] = srcs[k];
}

}  // namespace

void build_network(index_t n_total, std::uint64_t seed, Network& net) {
    if (n_total == 0) throw std::invalid_argument("N_total must be at least 1");
    if (n_total > kMaxNTotal) throw std::invalid_argument("N_total exceeds the 32-bit edge-id limit");
    const ScalingConfig cfg = make_scaling(n_total);
    net.cfg = cfg;
    net.seed = seed;
    const index_t n_exc = cfg.n_exc, n_inh = cfg.n_inh, n_astro = cfg.n_astro;
    const index_t n_neurons = n_exc + n_inh;
    const auto t0 = std::chrono::steady_clock::now();

    // ---- connectivity: one sequential setup stream (src/astrosim_fused.cu:631-735) ----
    HostRng rng(seed, kSetupStream);
    // Skip the draws the per-cell parameter generation would make first
    // (src/astrosim_fused.cu:644-645).
    for (index_t i = 0; i < n_astro; ++i) {
        rng.normal(); rng.normal(); rng.normal(); rng.normal();
    }
    for (index_t i = 0; i < n_neurons; ++i) {
        rng.normal(); rng.normal();
    }

    // Astrocyte pools, `take` distinct astrocytes per postsynaptic neuron
    // (src/astrosim_fused.cu:651-671).
    const int take = std::min<int>(ScalingConfig::kPoolSize, static_cast<int>(n_astro));
    std::vector<index_t> pool(std::size_t(n_neurons) * take);
    {
        std::vector<index_t> candidates;
        for (index_t post = 0; post < n_neurons; ++post) {
            index_t* pl = &pool[std::size_t(post) * take];
            if (static_cast<std::int64_t>(take) * 4 >= static_cast<std::int64_t>(n_astro)) {
                candidates.resize(n_astro);
                for (index_t i = 0; i < n_astro; ++i) candidates[i] = i;
                for (int k = 0; k < take; ++k) {
                    const auto pick = static_cast<std::size_t>(rng.uniform() * (candidates.size() - k)) + k;
                    std::swap(candidates[k], candidates[std::min(pick, candidates.size() - 1)]);
                    pl[k] = candidates[k];
                }
            } else {
                int have = 0;
                while (have < take) {
                    const auto cand =
                        static_cast<index_t>(rng.uniform() * static_cast<double>(n_astro)) % n_astro;
                    bool dup = false;
                    for (int q = 0; q < have; ++q) dup = dup || (pl[q] == cand);
                    if (!dup) pl[have++] = cand;
                }
            }
        }
    }

    const double p = cfg.p_primary;
    const double log1m_p = std::log1p(-p);
    std::vector<index_t> astro_out_degree(n_astro, 0);
    std::vector<index_t> indeg_exc(n_neurons, 0), indeg_inh(n_neurons, 0);
    SourceMajor exc_pass[2], inh_pass;
    std::vector<index_t> e2a_src, e2a_dst;  // (pre, astro) in generation order
    std::vector<index_t> sic_src, sic_dst;  // (astro, post) in generation order
    const std::size_t est = static_cast<std::size_t>(ScalingConfig::kSyn * 1.05 * n_neurons) + 64;
    exc_pass[0].dst.reserve(static_cast<std::size_t>(est * ScalingConfig::kEFraction));
    exc_pass[1].dst.reserve(static_cast<std::size_t>(est * (1.0 - ScalingConfig::kEFraction)) + 64);
    inh_pass.dst.reserve(static_cast<std::size_t>(est * (1.0 - ScalingConfig::kEFraction)) + 64);
    const std::size_t est3 = static_cast<std::size_t>(est * ScalingConfig::kPThird) + 64;
    e2a_src.reserve(est3); e2a_dst.reserve(est3);
    sic_src.reserve(est3); sic_dst.reserve(est3);

    // Excitatory rows: pre over [0, n_exc), posts over one block of targets,
    // with a third-factor astrocyte edge pair on 20% of synapses
    // (src/astrosim_fused.cu:688-710).
    auto connect_block = [&](SourceMajor& pass, index_t post_offset, index_t post_count) {
        pass.off.assign(std::size_t(n_exc) + 1, 0);
        for (index_t pre = 0; pre < n_exc; ++pre) {
            pass.off[pre] = pass.dst.size();
            index_t pos = 0;
            while (pos < post_count) {
                pos += geometric_skip(p, log1m_p, post_count - pos, rng);
                if (pos >= post_count) break;
                const index_t post = post_offset + pos;
                ++pos;
                if (pre == post) continue;  // no autapses
                pass.dst.push_back(post);
                indeg_exc[post]++;
                if (rng.uniform() >= ScalingConfig::kPThird) continue;  // n_astro >= 1 always
                const index_t* pl = &pool[std::size_t(post) * take];
                const index_t astro =
                    pl[static_cast<std::size_t>(rng.uniform() * take) % static_cast<std::size_t>(take)];
                if (astro_out_degree[astro] >= ScalingConfig::kMaxAstroOutDegree) continue;
                e2a_src.push_back(pre); e2a_dst.push_back(astro);
                sic_src.push_back(astro); sic_dst.push_back(post);
                ++astro_out_degree[astro];
            }
        }
        pass.off[n_exc] = pass.dst.size();
    };
    connect_block(exc_pass[0], 0, n_exc);
    connect_block(exc_pass[1], n_exc, n_inh);

    // Inhibitory rows; the source id kept is the LOCAL inhibitory index
    // (src/astrosim_fused.cu:712-723).
    inh_pass.off.assign(std::size_t(n_inh) + 1, 0);
    for (index_t pre = 0; pre < n_inh; ++pre) {
        inh_pass.off[pre] = inh_pass.dst.size();
        const index_t pre_global = n_exc + pre;
        index_t pos = 0;
        while (pos < n_neurons) {
            pos += geometric_skip(p, log1m_p, n_neurons - pos, rng);
            if (pos >= n_neurons) break;
            const index_t post = pos;
            ++pos;
            if (pre_global == post) continue;
            inh_pass.dst.push_back(post);
            indeg_inh[post]++;
        }
    }
    inh_pass.off[n_inh] = inh_pass.dst.size();

    // Four stable inversions, independent of each other: run them in parallel.
    {
        std::thread t_exc([&] { invert(n_neurons, {&exc_pass[0], &exc_pass[1]}, indeg_exc, net.inc_exc); });
        std::thread t_inh([&] { invert(n_neurons, {&inh_pass}, indeg_inh, net.inc_inh); });
        std::thread t_sic([&] { invert_pairs(n_neurons, sic_src, sic_dst, net.inc_sic); });
        invert_pairs(n_astro, e2a_src, e2a_dst, net.inc_e2a);
        t_exc.join();
        t_inh.join();
        t_sic.join();
    }
    const auto t1 = std::chrono::steady_clock::now();
    net.build_seconds = std::chrono::duration<double>(t1 - t0).count();

    // ---- per-cell parameters: a fresh stream from counter 0 (src/astrosim_fused.cu:917-933) ----
    const Biology bio;
    HostRng prng(seed, kSetupStream);
    net.Ca_tot.resize(n_astro); net.IP3_0.resize(n_astro);
    net.tau_IP3.resize(n_astro); net.delta_IP3.resize(n_astro);
    for (index_t i = 0; i < n_astro; ++i) {
        net.Ca_tot[i] = draw_scaled(bio.Ca_tot, bio.randomize_astro, bio.ra_lower, bio.ra_upper, bio.ra_var, prng);
        net.IP3_0[i] = draw_scaled(bio.IP3_0, bio.randomize_astro, bio.ra_lower, bio.ra_upper, bio.ra_var, prng);
        net.tau_IP3[i] = draw_scaled(bio.tau_IP3, bio.randomize_astro, bio.ra_lower, bio.ra_upper, bio.ra_var, prng);
        net.delta_IP3[i] = draw_scaled(bio.delta_IP3, bio.randomize_astro, bio.ra_lower, bio.ra_upper, bio.ra_var, prng);
    }
    net.V_reset.resize(n_neurons); net.b.resize(n_neurons);
    for (index_t i = 0; i < n_neurons; ++i) {
        const bool exc = i < n_exc;
        net.V_reset[i] = draw_scaled(exc ? bio.e_V_reset : bio.i_V_reset, bio.randomize_neuron,
                                     bio.rn_lower, bio.rn_upper, bio.rn_var, prng);
        net.b[i] = draw_scaled(exc ? bio.e_b : bio.i_b, bio.randomize_neuron, bio.rn_lower,
                               bio.rn_upper, bio.rn_var, prng);
    }
}

bool dump_edges_csv(const Network& net, const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return false;
    std::fputs("type,source,target\n", f);
    auto dump = [&](const char* type, const HostCSR& csr, long long src_offset, const char* sp,
                    const char* tp) {
        for (index_t t = 0; t < csr.n_targets(); ++t) {
            for (index_t k = csr.row_start[t]; k < csr.row_start[t + 1]; ++k) {
                std::fprintf(f, "%s,%s%lld,%s%u\n", type, sp,
                             static_cast<long long>(csr.src[k]) + src_offset, tp, t);
            }
        }
    };
    dump("exc_primary", net.inc_exc, 0, "", "");
    dump("inh_primary", net.inc_inh, static_cast<long long>(net.cfg.n_exc), "", "");
    dump("n2a", net.inc_e2a, 0, "", "A");
    dump("a2n", net.inc_sic, 0, "A", "");
    const bool ok = std::ferror(f) == 0;
    return (std::fclose(f) == 0) && ok;
}

}  // namespace astrosim
