// as_setup.cpp -- host network construction (see as_setup.h).
#include "as_setup.h"

#include "as_gpubuild.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <future>
#include <memory>
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
    c.ratio_ER_cyt_div = make_fixed_divisor(c.ratio_ER_cyt);
    return c;
}

NeuronConst make_neuron_const(const Biology& bio, bool exc, double h_step) {
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
    // the spike-current exponent (V - V_th) / Delta_T is at most this (V is
    // clamped at V_peak), which the FP64 step's exp requires to stay <= 709
    if (!(p.Delta_T > 0.0) || !((p.V_peak - p.V_th) / p.Delta_T <= 709.0))
        throw std::invalid_argument("AdEx constants out of range: (V_peak - V_th) / Delta_T must be in (-inf, 709]");
    p.Delta_T_div = make_fixed_divisor(p.Delta_T);
    p.C_m_div = make_fixed_divisor(p.C_m);
    p.tau_w_div = make_fixed_divisor(p.tau_w);
    p.refractory_steps = static_cast<int>(p.t_ref / h_step + 0.5);
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

void HostRng::skip_normals(std::uint64_t count) {
    // Each pair of normal() calls draws u1 (redrawn while u1 <= 1e-300) and
    // u2 once; only the counter matters to later uniform() draws, and the
    // transcendentals are evaluated solely for a trailing unpaired call.
    while (count > 0) {
        if (has_spare_) {
            has_spare_ = false;
            --count;
            continue;
        }
        if (count == 1) {
            normal();
            return;
        }
        while (uniform() <= 1e-300) {
        }
        ++counter_;  // u2
        count -= 2;
    }
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

unsigned worker_count() {
    const unsigned hc = std::thread::hardware_concurrency();
    return std::max(1u, std::min(hc == 0 ? 4u : hc, 16u));
}

// Runs f(t) for t in [0, n) on n threads (the calling thread takes t = 0).
template <class F>
void parallel_for_threads(unsigned n, F&& f) {
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < n; ++t) pool.emplace_back([&f, t] { f(t); });
    f(0);
    for (auto& th : pool) th.join();
}

// Geometric-skip values are a pure function of the setup stream's counter:
// G(c) = floor(log1p(-u_c) / log1p(-p)) with u_c = rng_uniform(seed, stream, c),
// saturated at 2^32-1 (every `remaining` fits in 32 bits, so the saturated
// value gives the same min(G, remaining)). Worker threads fill the values
// for the next window of counters while the sequential generator consumes
// the current one; the generator takes a value only where the reference
// draws a skip, so the draw sequence is exactly the reference's.
class SkipWindows {
public:
    // window: counters per window (sized to the network so small builds do
    // not precompute far past their last draw). take: astrocyte pool size.
    SkipWindows(std::uint64_t seed, std::uint64_t stream, double log1m_p, int take, std::uint64_t first,
                unsigned threads, std::uint64_t window)
        : kWindow(window), seed_(seed), stream_(stream), log1m_p_(log1m_p), take_(take), threads_(threads) {
        cur_.resize(kWindow);
        nxt_.resize(kWindow);
        cur_aux_.resize(kWindow);
        nxt_aux_.resize(kWindow);
        cur_base_ = first;
        fill(cur_, cur_aux_, cur_base_);
        pending_ = std::async(std::launch::async, [this] { fill(nxt_, nxt_aux_, cur_base_ + kWindow); });
    }
    ~SkipWindows() {
        if (pending_.valid()) pending_.wait();
    }

    // The draw at counter c read as a geometric skip.
    std::uint32_t skip(std::uint64_t c) { return cur_[slide(c)]; }
    // The draw at counter c read as the third-factor coin (bit 0) and as the
    // pool pick (bits 1-3).
    std::uint8_t aux(std::uint64_t c) { return cur_aux_[slide(c)]; }

private:
    std::size_t slide(std::uint64_t c) {
        if (c >= cur_base_ + kWindow) {
            pending_.wait();
            std::swap(cur_, nxt_);
            std::swap(cur_aux_, nxt_aux_);
            cur_base_ += kWindow;
            pending_ = std::async(std::launch::async, [this] { fill(nxt_, nxt_aux_, cur_base_ + kWindow); });
        }
        return static_cast<std::size_t>(c - cur_base_);
    }

    // g = floor(log1p(-u) / log1p(-p)) with glibc's log1p, as the reference.
    // The fast evaluation (gb::skip_fast) decides unless the quotient lies
    // too close to an integer; then glibc's log1p decides.
    static std::uint32_t skip_of(double u, double log1m_p) {
        std::uint32_t g;
        if (gb::skip_fast(u, log1m_p, g)) return g;
        const double q = std::floor(std::log1p(-u) / log1m_p);
        return q >= 4294967295.0 ? 0xFFFFFFFFu : static_cast<std::uint32_t>(q);
    }

    // Every counter is decoded all three ways the walk may read it; the
    // expressions are those of the sequential draws (src/astrosim_fused.cu:
    // 617-624 for the skip, 703-706 for the coin and the pool pick).
    void fill(std::vector<std::uint32_t>& buf, std::vector<std::uint8_t>& aux, std::uint64_t base) {
        parallel_for_threads(threads_, [&](unsigned t) {
            const std::uint64_t lo = kWindow * t / threads_, hi = kWindow * (t + 1) / threads_;
            for (std::uint64_t i = lo; i < hi; ++i) {
                const double r = rng_uniform(seed_, stream_, base + i);
                const double u = r >= 1.0 ? std::nextafter(1.0, 0.0) : r;
                buf[i] = skip_of(u, log1m_p_);
                const unsigned coin = r < ScalingConfig::kPThird ? 1u : 0u;
                const unsigned slot = static_cast<unsigned>(static_cast<std::size_t>(r * take_) %
                                                            static_cast<std::size_t>(take_));
                aux[i] = static_cast<std::uint8_t>(coin | (slot << 1));
            }
        });
    }

    const std::uint64_t kWindow;
    std::uint64_t seed_, stream_;
    double log1m_p_;
    int take_;
    unsigned threads_;
    std::vector<std::uint32_t> cur_, nxt_;
    std::vector<std::uint8_t> cur_aux_, nxt_aux_;
    std::uint64_t cur_base_ = 0;
    std::future<void> pending_;
};

// Window-fill threads beside the sequential walk: leave the walk a hardware
// thread of its own when the build would otherwise occupy all of them.
unsigned fill_threads(unsigned threads) {
    const unsigned hc = std::max(1u, std::thread::hardware_concurrency());
    return std::max(1u, std::min(threads, hc > 1 ? hc - 1 : 1u));
}

// Sequential draws on the setup stream, decoded from SkipWindows (or computed
// directly when p is 0 or 1). Mirrors HostRng::uniform and geometric_skip
// (src/astrosim_fused.cu:183, 617-624) draw for draw.
class SetupDraws {
public:
    SetupDraws(std::uint64_t seed, std::uint64_t counter, double p, int take, unsigned threads,
               std::uint64_t expected_draws)
        : seed_(seed), counter_(counter), p_(p), take_(take) {
        std::uint64_t window = std::uint64_t(1) << 16;
        while (window < (std::uint64_t(1) << 22) && window * 8 < expected_draws) window <<= 1;
        if (p > 0.0 && p < 1.0)
            skips_.reset(new SkipWindows(seed, kSetupStream, std::log1p(-p), take, counter, fill_threads(threads),
                                         window));
    }
    index_t skip(index_t remaining) {
        if (remaining == 0 || p_ >= 1.0) return 0;
        if (p_ <= 0.0) return remaining;
        const std::uint32_t g = skips_->skip(counter_++);
        return g >= remaining ? remaining : g;
    }
    // uniform() < kPThird
    bool third_factor() {
        if (skips_) return (skips_->aux(counter_++) & 1u) != 0;
        return uniform() < ScalingConfig::kPThird;
    }
    // static_cast<size_t>(uniform() * take) % take
    std::size_t pool_pick() {
        if (skips_) return skips_->aux(counter_++) >> 1;
        return static_cast<std::size_t>(uniform() * take_) % static_cast<std::size_t>(take_);
    }

private:
    double uniform() { return rng_uniform(seed_, kSetupStream, counter_++); }

    std::uint64_t seed_, counter_;
    double p_;
    int take_;
    std::unique_ptr<SkipWindows> skips_;
};

// Stable inversion of a generation-ordered edge sequence into a
// target-major CSR, as the reference's counting sort (count, prefix,
// scatter in generation order; src/astrosim_fused.cu:307-314), in O(edges)
// work over the threads. `visit(lo, hi, emit)` calls emit(src, dst) for the
// sequence positions [lo, hi) in order.
//
// Two streaming levels keep the writes cache-friendly: the sequence, cut into
// one contiguous chunk per thread, is first split by target range (a few
// thousand targets per range) into per-range lists, chunks in order; then
// each range is counting-sorted on its own, with its counters and its rows of
// the output in cache. Both levels keep generation order, so every row lists
// its sources exactly as the reference's counting sort does.
template <class Visit>
void invert_stable(index_t n_targets, std::size_t n_edges, unsigned threads, Visit&& visit, HostCSR& out) {
    out.row_start.assign(std::size_t(n_targets) + 1, 0);
    out.src.resize(n_edges);  // every element is written below
    if (n_targets == 0 || n_edges == 0) return;
    // ranges of 2^shift targets, at most 4096 of them
    int shift = 12;
    while ((std::uint64_t(n_targets) >> shift) >= 4096) ++shift;
    const std::size_t R = (std::size_t(n_targets) + (std::size_t(1) << shift) - 1) >> shift;
    const unsigned T = static_cast<unsigned>(
        std::max<std::uint64_t>(1, std::min<std::uint64_t>(threads, std::max<std::uint64_t>(1, n_edges >> 16))));
    auto chunk_lo = [&](unsigned t) { return static_cast<std::size_t>(std::uint64_t(n_edges) * t / T); };
    // level 1: per (chunk, range) counts, then offsets with chunks in order inside each range
    std::vector<std::uint64_t> off(std::size_t(T) * R + 1, 0);  // [range][chunk]
    parallel_for_threads(T, [&](unsigned t) {
        std::vector<std::uint64_t> c(R, 0);
        visit(chunk_lo(t), chunk_lo(t + 1), [&](index_t, index_t d) { ++c[d >> shift]; });
        for (std::size_t r = 0; r < R; ++r) off[r * T + t] = c[r];
    });
    {
        std::uint64_t run = 0;
        for (std::size_t i = 0; i < std::size_t(T) * R; ++i) {
            const std::uint64_t c = off[i];
            off[i] = run;
            run += c;
        }
        off[std::size_t(T) * R] = run;
    }
    struct Pair { index_t src, dst; };
    std::unique_ptr<Pair[]> pairs(new Pair[n_edges]);  // grouped by range; first touched by the writers
    parallel_for_threads(T, [&](unsigned t) {
        std::vector<std::uint64_t> cur(R);
        for (std::size_t r = 0; r < R; ++r) cur[r] = off[r * T + t];
        Pair* p = pairs.get();
        visit(chunk_lo(t), chunk_lo(t + 1), [&](index_t sv, index_t d) { p[cur[d >> shift]++] = Pair{sv, d}; });
    });
    // level 2: each range's in-degrees, the row starts, then a stable scatter
    std::vector<std::uint64_t> range_edges(R + 1, 0);
    for (std::size_t r = 0; r < R; ++r) range_edges[r + 1] = off[(r + 1) * T];
    std::atomic<std::size_t> next{0};
    parallel_for_threads(T, [&](unsigned) {
        std::vector<index_t> cur;
        for (std::size_t r; (r = next.fetch_add(1)) < R;) {
            const std::size_t d0 = r << shift;
            const std::size_t d1 = std::min<std::size_t>(std::size_t(n_targets), d0 + (std::size_t(1) << shift));
            const Pair* p0 = pairs.get() + range_edges[r];
            const Pair* p1 = pairs.get() + range_edges[r + 1];
            cur.assign(d1 - d0, 0);
            for (const Pair* q = p0; q != p1; ++q) ++cur[q->dst - d0];
            std::uint64_t run = range_edges[r];
            for (std::size_t i = 0; i < d1 - d0; ++i) {
                const index_t c = cur[i];
                cur[i] = static_cast<index_t>(run);
                run += c;
                out.row_start[d0 + i + 1] = static_cast<index_t>(run);
            }
            index_t* src = out.src.data();
            for (const Pair* q = p0; q != p1; ++q) src[cur[q->dst - d0]++] = q->src;
        }
    });
}

// Edges of one source-major generation pass: the targets of source s are
// dst[off[s] .. off[s+1]), in generation order.
struct SourceMajor {
    std::vector<index_t> dst;
    std::vector<std::uint64_t> off;
};

}  // namespace

std::vector<index_t> make_pools(const ScalingConfig& cfg, HostRng& rng) {
    // Astrocyte pools, `take` distinct astrocytes per postsynaptic neuron
    // (src/astrosim_fused.cu:651-671).
    const index_t n_astro = cfg.n_astro, n_neurons = cfg.n_neurons();
    const int take = std::min<int>(ScalingConfig::kPoolSize, static_cast<int>(n_astro));
    std::vector<index_t> pool(std::size_t(n_neurons) * take);
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
                const auto cand = static_cast<index_t>(rng.uniform() * static_cast<double>(n_astro)) % n_astro;
                bool dup = false;
                for (int q = 0; q < have; ++q) dup = dup || (pl[q] == cand);
                if (!dup) pl[have++] = cand;
            }
        }
    }
    return pool;
}

namespace {

// The Gaussian draws of a fresh setup stream, computed window by window on
// worker threads. normal() number k of HostRng from counter 0 is the cosine
// (k even) or the sine (k odd) of Box-Muller pair k/2 on counters k/2*2 and
// k/2*2+1, unless an earlier pair redrew its u1 (u1 <= 1e-300, probability
// 2^-53 per pair): a window holding such a pair reports it, and the caller
// then draws sequentially with HostRng.
class NormalWindows {
public:
    NormalWindows(std::uint64_t seed, unsigned threads, std::uint64_t expected)
        : seed_(seed), threads_(threads), expected_(expected) {}
    bool normal(double& z) {
        if (pos_ == buf_.size() && !fill()) return false;
        z = buf_[pos_++];
        return true;
    }

private:
    bool fill() {
        // the first window covers the expected draws, later ones a little more
        const std::uint64_t pairs = next_pair_ == 0 ? expected_ / 2 + 1024
                                                    : std::max<std::uint64_t>(4096, expected_ / 64);
        buf_.resize(2 * pairs);
        std::atomic<bool> redrawn{false};
        const unsigned T = static_cast<unsigned>(std::max<std::uint64_t>(1, std::min<std::uint64_t>(threads_, pairs / 4096)));
        parallel_for_threads(T, [&](unsigned t) {
            for (std::uint64_t i = pairs * t / T; i < pairs * (t + 1) / T; ++i) {
                const std::uint64_t c = 2 * (next_pair_ + i);
                const double u1 = rng_uniform(seed_, kSetupStream, c);
                if (u1 <= 1e-300) {
                    redrawn = true;
                    continue;
                }
                const double u2 = rng_uniform(seed_, kSetupStream, c + 1);
                // the expressions of HostRng::normal
                const double r = std::sqrt(-2.0 * std::log(u1));
                const double theta = 6.283185307179586476925286766559 * u2;
                double sn, cs;
                ::sincos(theta, &sn, &cs);
                buf_[2 * i] = r * cs;
                buf_[2 * i + 1] = r * sn;
            }
        });
        if (redrawn) return false;
        next_pair_ += pairs;
        pos_ = 0;
        return true;
    }
    std::uint64_t seed_;
    unsigned threads_;
    std::uint64_t expected_, next_pair_ = 0;
    std::vector<double> buf_;
    std::size_t pos_ = 0;
};

// draw_scaled over a source of normals `nrm` (bool(double&); false: the
// source gave out): the same bounds, tries and clamp as HostRng's
// normal_redraw (src/astrosim_fused.cu:196-216).
template <class Nrm>
bool scaled_from(double mean, bool enabled, double lower, double upper, double var, Nrm& nrm, double& out) {
    if (!enabled || var <= 0.0) {
        out = mean;
        return true;
    }
    const double lo = mean < 0.0 ? mean * upper : mean * lower;
    const double hi = mean < 0.0 ? mean * lower : mean * upper;
    const double stddev = var * std::fabs(mean);
    for (int i = 0; i < 1000; ++i) {
        double z;
        if (!nrm(z)) return false;
        const double v = mean + stddev * z;
        if (v >= lo && v <= hi) {
            out = v;
            return true;
        }
    }
    out = std::min(std::max(mean, lo), hi);
    return true;
}

// The per-cell draws in the reference's order (src/astrosim_fused.cu:917-933).
template <class Nrm>
bool cell_params_from(const ScalingConfig& cfg, Nrm& nrm, Network& net) {
    const Biology bio;
    const index_t n_astro = cfg.n_astro, n_neurons = cfg.n_neurons(), n_exc = cfg.n_exc;
    net.Ca_tot.resize(n_astro); net.IP3_0.resize(n_astro);
    net.tau_IP3.resize(n_astro); net.delta_IP3.resize(n_astro);
    for (index_t i = 0; i < n_astro; ++i) {
        if (!scaled_from(bio.Ca_tot, bio.randomize_astro, bio.ra_lower, bio.ra_upper, bio.ra_var, nrm, net.Ca_tot[i]) ||
            !scaled_from(bio.IP3_0, bio.randomize_astro, bio.ra_lower, bio.ra_upper, bio.ra_var, nrm, net.IP3_0[i]) ||
            !scaled_from(bio.tau_IP3, bio.randomize_astro, bio.ra_lower, bio.ra_upper, bio.ra_var, nrm, net.tau_IP3[i]) ||
            !scaled_from(bio.delta_IP3, bio.randomize_astro, bio.ra_lower, bio.ra_upper, bio.ra_var, nrm,
                         net.delta_IP3[i]))
            return false;
    }
    net.V_reset.resize(n_neurons); net.b.resize(n_neurons);
    for (index_t i = 0; i < n_neurons; ++i) {
        const bool exc = i < n_exc;
        if (!scaled_from(exc ? bio.e_V_reset : bio.i_V_reset, bio.randomize_neuron, bio.rn_lower, bio.rn_upper,
                         bio.rn_var, nrm, net.V_reset[i]) ||
            !scaled_from(exc ? bio.e_b : bio.i_b, bio.randomize_neuron, bio.rn_lower, bio.rn_upper, bio.rn_var, nrm,
                         net.b[i]))
            return false;
    }
    return true;
}

}  // namespace

void draw_cell_params(const ScalingConfig& cfg, std::uint64_t seed, Network& net) {
    // A fresh stream from counter 0 (src/astrosim_fused.cu:917-933). About
    // 1.88 normals per astrocyte parameter (bounds [0.9, 1.05] x mean at 10%)
    // and 1.05 per neuron parameter ([0.9, 1.1] at 5%).
    const std::uint64_t expected =
        static_cast<std::uint64_t>(4.0 * 1.876 * cfg.n_astro + 2.0 * 1.048 * cfg.n_neurons()) + 64;
    const unsigned threads = std::max(1u, std::min(worker_count(), static_cast<unsigned>(cfg.n_total / 65536u) + 1u));
    NormalWindows win(seed, threads, expected);
    auto from_windows = [&](double& z) { return win.normal(z); };
    if (cell_params_from(cfg, from_windows, net)) return;
    HostRng prng(seed, kSetupStream);
    auto sequential = [&](double& z) {
        z = prng.normal();
        return true;
    };
    cell_params_from(cfg, sequential, net);
}

void build_network(index_t n_total, std::uint64_t seed, Network& net) {
    if (n_total == 0) throw std::invalid_argument("N_total must be at least 1");
    if (n_total > kMaxNTotal) throw std::invalid_argument("N_total exceeds the 32-bit edge-id limit");
    const ScalingConfig cfg = make_scaling(n_total);
    net.cfg = cfg;
    net.seed = seed;
    const index_t n_exc = cfg.n_exc, n_inh = cfg.n_inh, n_astro = cfg.n_astro;
    const index_t n_neurons = n_exc + n_inh;
    // threads for this build: small networks are usually built many at a time
    const unsigned threads = std::max(1u, std::min(worker_count(), static_cast<unsigned>(n_total / 65536u) + 1u));
    const auto t0 = std::chrono::steady_clock::now();

    // ---- connectivity: one sequential setup stream (src/astrosim_fused.cu:631-735) ----
    HostRng rng(seed, kSetupStream);
    // Skip the draws the per-cell parameter generation would make first
    // (src/astrosim_fused.cu:644-645).
    rng.skip_normals(4ull * n_astro + 2ull * n_neurons);

    const int take = std::min<int>(ScalingConfig::kPoolSize, static_cast<int>(n_astro));
    std::vector<index_t> pool = make_pools(cfg, rng);
    const auto t_pools = std::chrono::steady_clock::now();

    // Edge generation. Excitatory rows: pre over [0, n_exc), posts over one
    // block of targets, a third-factor astrocyte pick on 20% of synapses
    // (src/astrosim_fused.cu:688-710); inhibitory rows keep the LOCAL source
    // id (src/astrosim_fused.cu:712-723). Third-factor synapses are recorded
    // as (pre, post, pool slot) and resolved after the sequential pass.
    const double p = cfg.p_primary;
    // about 2.3 draws per synapse (skip, third-factor coin, pool pick on 20%)
    SetupDraws dr(seed, rng.counter(), p, take, threads,
                  static_cast<std::uint64_t>(2.5 * ScalingConfig::kSyn * n_neurons));
    SourceMajor exc_pass[2], inh_pass;
    std::vector<index_t> tf_pre, tf_post;
    std::vector<std::uint8_t> tf_slot;
    const std::size_t est = static_cast<std::size_t>(ScalingConfig::kSyn * 1.02 * n_neurons) + 64;
    exc_pass[0].dst.reserve(static_cast<std::size_t>(est * ScalingConfig::kEFraction));
    exc_pass[1].dst.reserve(static_cast<std::size_t>(est * (1.0 - ScalingConfig::kEFraction)) + 64);
    inh_pass.dst.reserve(static_cast<std::size_t>(est * (1.0 - ScalingConfig::kEFraction)) + 64);
    const std::size_t est3 = static_cast<std::size_t>(est * ScalingConfig::kPThird * 1.02) + 64;
    tf_pre.reserve(est3); tf_post.reserve(est3); tf_slot.reserve(est3);

    auto connect_block = [&](SourceMajor& pass, index_t post_offset, index_t post_count) {
        pass.off.assign(std::size_t(n_exc) + 1, 0);
        for (index_t pre = 0; pre < n_exc; ++pre) {
            pass.off[pre] = pass.dst.size();
            index_t pos = 0;
            while (pos < post_count) {
                pos += dr.skip(post_count - pos);
                if (pos >= post_count) break;
                const index_t post = post_offset + pos;
                ++pos;
                if (pre == post) continue;  // no autapses
                pass.dst.push_back(post);
                if (!dr.third_factor()) continue;  // n_astro >= 1 always
                const std::size_t slot = dr.pool_pick();
                tf_pre.push_back(pre);
                tf_post.push_back(post);
                tf_slot.push_back(static_cast<std::uint8_t>(slot));
            }
        }
        pass.off[n_exc] = pass.dst.size();
    };
    connect_block(exc_pass[0], 0, n_exc);
    connect_block(exc_pass[1], n_exc, n_inh);
    inh_pass.off.assign(std::size_t(n_inh) + 1, 0);
    for (index_t pre = 0; pre < n_inh; ++pre) {
        inh_pass.off[pre] = inh_pass.dst.size();
        const index_t pre_global = n_exc + pre;
        index_t pos = 0;
        while (pos < n_neurons) {
            pos += dr.skip(n_neurons - pos);
            if (pos >= n_neurons) break;
            const index_t post = pos;
            ++pos;
            if (pre_global == post) continue;
            inh_pass.dst.push_back(post);
        }
    }
    inh_pass.off[n_inh] = inh_pass.dst.size();

    // Third-factor synapses: astrocyte = pool[post][slot] (parallel), then the
    // out-degree cap applied in generation order (sequential, as the reference).
    const std::size_t n_tf = tf_pre.size();
    std::vector<index_t> tf_astro(n_tf);
    parallel_for_threads(threads, [&](unsigned t) {
        const std::size_t lo = n_tf * t / threads, hi = n_tf * (t + 1) / threads;
        for (std::size_t i = lo; i < hi; ++i) tf_astro[i] = pool[std::size_t(tf_post[i]) * take + tf_slot[i]];
    });
    std::vector<std::uint8_t>().swap(tf_slot);
    std::vector<index_t>().swap(pool);
    std::vector<index_t> astro_out_degree(n_astro, 0);
    std::size_t kept = 0;
    for (std::size_t i = 0; i < n_tf; ++i) {
        const index_t a = tf_astro[i];
        if (astro_out_degree[a] >= ScalingConfig::kMaxAstroOutDegree) continue;
        ++astro_out_degree[a];
        tf_pre[kept] = tf_pre[i];
        tf_post[kept] = tf_post[i];
        tf_astro[kept] = a;
        ++kept;
    }
    tf_pre.resize(kept); tf_post.resize(kept); tf_astro.resize(kept);
    const auto t_gen = std::chrono::steady_clock::now();

    // Four stable inversions over the worker threads.
    // Sequence positions of a source-major pass: source s owns [off[s], off[s+1]).
    auto visit_pass = [](const SourceMajor& pass, std::uint64_t lo, std::uint64_t hi, auto&& emit) {
        if (lo >= hi) return;
        std::size_t s = static_cast<std::size_t>(
            std::upper_bound(pass.off.begin(), pass.off.end(), lo) - pass.off.begin()) - 1;
        for (std::uint64_t k = lo; k < hi; ++k) {
            while (pass.off[s + 1] <= k) ++s;
            emit(static_cast<index_t>(s), pass.dst[k]);
        }
    };
    {
        // pass 0 (targets < n_exc) then pass 1 (targets >= n_exc): no target in both
        const std::uint64_t e0 = exc_pass[0].dst.size(), e1 = exc_pass[1].dst.size();
        invert_stable(n_neurons, e0 + e1, threads, [&](std::uint64_t lo, std::uint64_t hi, auto&& emit) {
            visit_pass(exc_pass[0], lo, std::min(hi, e0), emit);
            visit_pass(exc_pass[1], std::max(lo, e0) - e0, std::max(hi, e0) - e0, emit);
        }, net.inc_exc);
    }
    exc_pass[0] = SourceMajor();
    exc_pass[1] = SourceMajor();
    invert_stable(n_neurons, inh_pass.dst.size(), threads, [&](std::uint64_t lo, std::uint64_t hi, auto&& emit) {
        visit_pass(inh_pass, lo, hi, emit);
    }, net.inc_inh);
    inh_pass = SourceMajor();
    invert_stable(n_neurons, kept, threads, [&](std::uint64_t lo, std::uint64_t hi, auto&& emit) {
        for (std::uint64_t i = lo; i < hi; ++i) emit(tf_astro[i], tf_post[i]);
    }, net.inc_sic);
    invert_stable(n_astro, kept, threads, [&](std::uint64_t lo, std::uint64_t hi, auto&& emit) {
        for (std::uint64_t i = lo; i < hi; ++i) emit(tf_pre[i], tf_astro[i]);
    }, net.inc_e2a);
    const auto t1 = std::chrono::steady_clock::now();
    net.build_seconds = std::chrono::duration<double>(t1 - t0).count();
    net.phase_seconds[0] = std::chrono::duration<double>(t_pools - t0).count();
    net.phase_seconds[1] = std::chrono::duration<double>(t_gen - t_pools).count();
    net.phase_seconds[2] = std::chrono::duration<double>(t1 - t_gen).count();

    // ---- per-cell parameters ----
    draw_cell_params(cfg, seed, net);
    net.phase_seconds[3] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
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
