// as_blocked.hip -- the blocked, event-driven FP64 method.
//
// Every synaptic delay is the same d >= 1 steps (1 ms at dt = 0.1 ms), so
// within an aligned block of L steps (L divides d, L <= 32) no cell reads
// anything written in the same block: all inputs of block b come from block
// b - d/L. One INTEGRATE launch therefore advances every astrocyte and
// neuron a whole block with its state in registers.
//
// Inputs are delivered event-driven instead of by probing every in-edge:
//   * each cell records the steps of its block at which it emitted (a spike,
//     or a nonzero SIC factor) as a bit word, the emitted values in a ring
//     indexed by absolute step, and appends itself to its block's emitter
//     list when the word is nonzero;
//   * a SCATTER launch walks only the emitters' out-edges (a transposed copy
//     of the in-CSR built on the device) and appends (in-CSR slot, source)
//     to the target's fixed-capacity bucket for that connection type;
//   * INTEGRATE sums a target's bucket in ascending in-CSR slot order, which
//     is the reference's CSR-order sum with its exact-zero terms removed
//     (src/astrosim_fused.cu:384-389, 450-463, 482-490), so FP64 results are
//     bit-identical to the plain method and to the host oracle. A bucket that
//     overflowed falls back to walking the whole in-row.
// Per block the work is: cells x L steps of integration (FP64 ALU bound)
// plus emitters x out-degree bucket appends.
#include <hipcub/hipcub.hpp>

#include "as_engine.h"
#include "as_model.h"
#include "as_model32.h"

namespace astrosim {
namespace {

constexpr int kSpikeCounters = 64;
constexpr unsigned kBucketCap = 8;
enum { T_EXC = 0, T_INH = 1, T_SIC = 2, T_E2A = 3 };

// Indices, buckets, emission words, emitter lists and counters: everything
// that does not depend on the state precision.
struct BlkCommon {
    // in-CSR by target (exc, inh, sic, e2a) and the transposed out-CSR by
    // source with (target, in-CSR slot) entries
    const index_t* in_rs[4];
    const index_t* in_src[4];
    const index_t* out_rs[4];
    const uint2* out_ent[4];
    index_t out_nsrc[4];
    // buckets per type: count per target, kBucketCap (slot, source) entries
    unsigned* bcnt[4];
    uint2* bent[4];
    std::uint8_t* ever;
    std::int32_t* refr;
    std::uint32_t* spk_w;  // [H][n_neurons] spike bits of a block
    std::uint32_t* sic_w;  // [H][n_astro] nonzero-SIC bits of a block
    std::uint32_t* nlist;  // [H][n_neurons] neurons with a nonzero block word
    std::uint32_t* alist;  // [H][n_astro]
    std::uint32_t* lcount; // [H][2]
    unsigned long long* counters;
    unsigned* overflow;  // bucket overflows seen (diagnostic)
    index_t n_astro, n_neurons, n_exc, n_inh;
    std::uint32_t astro_blocks;
};

// State arrays in the simulation's precision R (double: FP64; float: FP32c,
// where Ca/IP3/h carry a compensation term and tau_IP3 holds 1/tau_IP3).
template <class R>
struct BlkState {
    R *Ca, *IP3, *h;
    R *Ca_lo, *IP3_lo, *h_lo;
    const R *Ca_tot, *IP3_0, *tau_IP3, *delta_IP3;
    R *V, *w, *g_ex, *dg_ex, *g_in, *dg_in, *I_sic;
    const R *V_reset, *b;
    R *stp_x, *stp_u, *stp_t;
    R* spk_v;  // [Rv][n_neurons] x*u at the spike step
    R* sic_v;  // [Rv][n_astro] SIC factor where nonzero
};

struct BlkLaunch {
    std::int64_t t0, bstart;
    int n;           // steps in this launch
    int d, Rv;       // delay (steps), value-ring depth
    std::uint32_t wslot, rslot;  // history slot written / read (block b, block b - d/L)
    int first, last;  // launch starts / completes its block
};

__device__ __forceinline__ unsigned lane_id() { return threadIdx.x & 63u; }

// Appends `id` (when `want`) to a list with one atomic per wave.
__device__ __forceinline__ void wave_append(bool want, std::uint32_t id, std::uint32_t* list,
                                            std::uint32_t* count) {
    const unsigned long long m = __ballot(want);
    if (m == 0ull) return;
    const unsigned leader = __ffsll((long long)m) - 1;
    std::uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(count, (std::uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (want) list[base + __popcll(m & ((1ull << lane_id()) - 1ull))] = id;
}

// Delayed-input sum of one target for block step k: the reference's in-row
// sum restricted to its nonzero terms, in ascending in-CSR slot order.
// `word(src)` is the source's emission bits for the read block and
// `val(src)` its emitted value at the delayed step.
template <class R, class WordF, class ValF>
__device__ R input_sum(const uint2* ent, unsigned cnt, const index_t* rs, const index_t* src,
                       index_t target, unsigned k, WordF word, ValF val) {
    R s = R(0);
    if (cnt > kBucketCap) {  // overflow: walk the whole row
        for (index_t e = rs[target]; e < rs[target + 1]; ++e) {
            const index_t q = src[e];
            if ((word(q) >> k) & 1u) s += val(q);
        }
        return s;
    }
    unsigned last = 0;
    for (unsigned r = 0; r < cnt; ++r) {
        unsigned best = 0xFFFFFFFFu, bsrc = 0;
        for (unsigned q = 0; q < cnt; ++q) {
            const uint2 x = ent[q];
            if ((r == 0 || x.x > last) && x.x < best) {
                best = x.x;
                bsrc = x.y;
            }
        }
        last = best;
        if ((word(bsrc) >> k) & 1u) s += val(bsrc);
    }
    return s;
}

// Union of the emission bits of a target's inputs (all ones on overflow).
template <class WordF>
__device__ std::uint32_t input_mask(const uint2* ent, unsigned cnt, WordF word) {
    if (cnt > kBucketCap) return 0xFFFFFFFFu;
    std::uint32_t m = 0;
    for (unsigned q = 0; q < cnt; ++q) m |= word(ent[q].y);
    return m;
}

// ---------------------------------------------------------------------------
// SCATTER: emitters of block b - d/L append to their targets' buckets.
// One wave per emitter, lanes over its out-edges.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void scatter_kernel(BlkCommon D, std::uint32_t rslot, std::uint32_t wslot) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        D.lcount[wslot * 2 + 0] = 0;
        D.lcount[wslot * 2 + 1] = 0;
    }
    const std::uint32_t nN = D.lcount[rslot * 2 + 0];
    const std::uint32_t nA = D.lcount[rslot * 2 + 1];
    const unsigned waves = gridDim.x * (blockDim.x / 64);
    const unsigned wave = blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    for (std::uint32_t i = wave; i < nN + nA; i += waves) {
        const bool is_neuron = i < nN;
        const std::uint32_t src = is_neuron ? D.nlist[std::size_t(rslot) * D.n_neurons + i]
                                            : D.alist[std::size_t(rslot) * D.n_astro + (i - nN)];
        // neuron: exc out-edges and n2a out-edges (sources < n_exc), and the
        // inhibitory rows whose stored LOCAL id equals this neuron id
        // (src/astrosim_fused.cu:459-462); astrocyte: its a2n out-edges.
        for (int pass = 0; pass < 3; ++pass) {
            int T;
            if (is_neuron) T = pass == 0 ? T_EXC : (pass == 1 ? T_INH : T_E2A);
            else {
                if (pass > 0) break;
                T = T_SIC;
            }
            if (src >= D.out_nsrc[T]) continue;
            const index_t lo = D.out_rs[T][src], hi = D.out_rs[T][src + 1];
            for (index_t e = lo + lane_id(); e < hi; e += 64) {
                const uint2 te = D.out_ent[T][e];  // (target, in-CSR slot)
                const unsigned slot = atomicAdd(&D.bcnt[T][te.x], 1u);
                if (slot < kBucketCap) D.bent[T][std::size_t(te.x) * kBucketCap + slot] = make_uint2(te.y, src);
                else if (slot == kBucketCap) atomicAdd(D.overflow, 1u);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// INTEGRATE: advance every cell over steps [t0, t0 + n) of one block.
// ---------------------------------------------------------------------------

// Value-ring row of the delayed step t - d (only dereferenced when a source's
// bit is set, which implies t >= d).
__device__ __forceinline__ std::size_t value_row(std::int64_t t, const BlkLaunch& B) {
    const std::int64_t tv = t - B.d;
    return tv >= 0 ? std::size_t(tv % B.Rv) : 0;
}

// One astrocyte over the launch's steps (src/astrosim_fused.cu:374-427).
template <class R>
__device__ bool astro_cell(const BlkCommon& D, const BlkState<R>& S, const RunParams& P, const Params32& Q,
                           std::uint64_t seed, const BlkLaunch& B, index_t a) {
    const std::size_t rN = std::size_t(B.rslot) * D.n_neurons;
    const unsigned cnt = D.bcnt[T_E2A][a];
    const uint2* ent = D.bent[T_E2A] + std::size_t(a) * kBucketCap;
    auto word = [&](index_t q) { return D.spk_w[rN + q]; };
    const std::uint32_t any = cnt ? input_mask(ent, cnt, word) : 0u;
    R ca = S.Ca[a], ip3 = S.IP3[a], hh = S.h[a];
    R ca_lo = R(0), ip3_lo = R(0), hh_lo = R(0);
    if constexpr (sizeof(R) == 4) {
        ca_lo = S.Ca_lo[a];
        ip3_lo = S.IP3_lo[a];
        hh_lo = S.h_lo[a];
    }
    const R Ca_tot = S.Ca_tot[a], IP3_0 = S.IP3_0[a], tau = S.tau_IP3[a], dIP3 = S.delta_IP3[a];
    std::uint32_t wout = B.first ? 0u : D.sic_w[std::size_t(B.wslot) * D.n_astro + a];
    bool ever = false;
    for (int i = 0; i < B.n; ++i) {
        const std::int64_t t = B.t0 + i;
        const unsigned k = static_cast<unsigned>(t - B.bstart);
        R gathered = R(0);
        if ((any >> k) & 1u) {
            const std::size_t vrow = value_row(t, B) * D.n_neurons;
            gathered = input_sum<R>(ent, cnt, D.in_rs[T_E2A], D.in_src[T_E2A], a, k, word,
                                    [&](index_t q) { return S.spk_v[vrow + q]; });
        }
        const std::uint64_t stream = std::uint64_t(t) * kStepStride + a;
        int ev = 0;
        if (P.poiss_astro) ev = rng_poisson(seed ^ kKeyAstroPoisson, stream, P.lam_astro, P.lim_astro);
        const double noise = P.noise_astro > 0.0 ? P.noise_astro * rng_normal(seed ^ kKeyAstroNoise, stream) : 0.0;
        R factor;
        if constexpr (sizeof(R) == 8) {
            double ip3_input = P.w_n2a * gathered;
            if (ev > 0) ip3_input += P.w_poiss_astro * static_cast<double>(ev);
            astro_advance(P.ac, Ca_tot, IP3_0, tau, dIP3, ip3_input, noise, P.h_step, P.substeps, ca, ip3, hh);
            factor = sic_factor(P.ac, ca);
        } else {
            float ip3_input = Q.w_n2a * gathered;
            if (ev > 0) ip3_input += Q.w_poiss_astro * static_cast<float>(ev);
            astro_advance32(Q.ac, Q, Ca_tot, IP3_0, tau, dIP3, ip3_input, static_cast<float>(noise), P.substeps, ca,
                            ca_lo, ip3, ip3_lo, hh, hh_lo);
            factor = sic_factor32(Q.ac, ca, ca_lo);
        }
        if (factor > R(0)) {
            ever = true;
            wout |= 1u << k;
            S.sic_v[std::size_t(t % B.Rv) * D.n_astro + a] = factor;
        }
    }
    S.Ca[a] = ca;
    S.IP3[a] = ip3;
    S.h[a] = hh;
    if constexpr (sizeof(R) == 4) {
        S.Ca_lo[a] = ca_lo;
        S.IP3_lo[a] = ip3_lo;
        S.h_lo[a] = hh_lo;
    }
    if (ever) D.ever[a] = 1;
    D.sic_w[std::size_t(B.wslot) * D.n_astro + a] = wout;
    if (B.last) D.bcnt[T_E2A][a] = 0;
    return B.last && wout != 0u;
}

// One neuron over the launch's steps (src/astrosim_fused.cu:433-587).
// Returns the number of spikes; *emit is set when the neuron joins the
// block's emitter list.
template <class R>
__device__ unsigned neuron_cell(const BlkCommon& D, const BlkState<R>& S, const RunParams& P, const Params32& Q,
                                std::uint64_t seed, const BlkLaunch& B, index_t j, bool* emit) {
    const std::size_t rN = std::size_t(B.rslot) * D.n_neurons;
    const std::size_t rA = std::size_t(B.rslot) * D.n_astro;
    const bool exc = j < D.n_exc;
    const unsigned ce = D.bcnt[T_EXC][j], ci = D.bcnt[T_INH][j], cs = D.bcnt[T_SIC][j];
    const uint2* ee = D.bent[T_EXC] + std::size_t(j) * kBucketCap;
    const uint2* ei = D.bent[T_INH] + std::size_t(j) * kBucketCap;
    const uint2* es = D.bent[T_SIC] + std::size_t(j) * kBucketCap;
    auto nword = [&](index_t q) { return D.spk_w[rN + q]; };
    auto aword = [&](index_t q) { return D.sic_w[rA + q]; };
    const std::uint32_t any_e = ce ? input_mask(ee, ce, nword) : 0u;
    const std::uint32_t any_i = ci ? input_mask(ei, ci, nword) : 0u;
    const std::uint32_t any_s = cs ? input_mask(es, cs, aword) : 0u;
    R V = S.V[j], w = S.w[j], g_ex = S.g_ex[j], dg_ex = S.dg_ex[j], g_in = S.g_in[j], dg_in = S.dg_in[j];
    int refr = D.refr[j];
    R I_sic = S.I_sic[j];
    const R V_reset = S.V_reset[j], bb = S.b[j];
    const double nstd = exc ? P.noise_exc : P.noise_inh;
    std::uint64_t noise_at = ~std::uint64_t(0);
    R noise_val = R(0);
    unsigned nspikes = 0;
    std::uint32_t wout = B.first ? 0u : D.spk_w[std::size_t(B.wslot) * D.n_neurons + j];
    for (int i = 0; i < B.n; ++i) {
        const std::int64_t t = B.t0 + i;
        const unsigned k = static_cast<unsigned>(t - B.bstart);
        const std::size_t vrow = value_row(t, B);
        auto nval = [&](index_t q) { return S.spk_v[vrow * D.n_neurons + q]; };
        R exc_sum = R(0), inh_sum = R(0);
        if ((any_e >> k) & 1u) exc_sum = input_sum<R>(ee, ce, D.in_rs[T_EXC], D.in_src[T_EXC], j, k, nword, nval);
        // inhibitory source ids are local indices used as neuron ids (src/astrosim_fused.cu:459-462)
        if ((any_i >> k) & 1u) inh_sum = input_sum<R>(ei, ci, D.in_rs[T_INH], D.in_src[T_INH], j, k, nword, nval);
        int ev = 0;
        if (exc ? P.poiss_exc : P.poiss_inh)
            ev = rng_poisson(seed ^ kKeyNeuronPoisson, std::uint64_t(t) * kStepStride + j, exc ? P.lam_exc : P.lam_inh,
                             exc ? P.lim_exc : P.lim_inh);
        if ((t - P.d_a2n >= 0) && ((t - P.d_a2n) % P.sic_interval == 0)) {
            R sic_sum = R(0);
            if ((any_s >> k) & 1u)
                sic_sum = input_sum<R>(es, cs, D.in_rs[T_SIC], D.in_src[T_SIC], j, k, aword,
                                       [&](index_t q) { return S.sic_v[vrow * D.n_astro + q]; });
            if constexpr (sizeof(R) == 8) I_sic = P.w_a2n * sic_sum;
            else I_sic = Q.w_a2n * sic_sum;
        }
        if (nstd > 0.0) {
            // the neuron noise stream changes every steps_per_noise steps
            const std::uint64_t nidx = static_cast<std::uint64_t>(t / P.steps_per_noise);
            if (nidx != noise_at) {
                noise_at = nidx;
                noise_val = static_cast<R>(nstd * rng_normal(seed ^ kKeyNeuronNoise, nidx * kNoiseStride + std::uint64_t(j)));
            }
        }
        bool spiked;
        if constexpr (sizeof(R) == 8) {
            const NeuronConst& p = exc ? P.ne : P.ni;
            double exc_input = P.w_e * exc_sum;
            const double inh_input = (-P.w_i) * inh_sum;
            if (ev > 0) exc_input += (exc ? P.w_poiss_exc : P.w_poiss_inh) * static_cast<double>(ev);
            double I_ext = I_sic;
            if (nstd > 0.0) I_ext += noise_val;
            NeuronRegs r{V, w, g_ex, dg_ex + exc_input * p.psc_init_ex, g_in, dg_in + inh_input * p.psc_init_in, refr};
            spiked = neuron_substeps(p, r, I_ext, V_reset, bb, P.h_step, P.substeps,
                                     exc ? P.ex_decay_e : P.ex_decay_i, exc ? P.in_decay_e : P.in_decay_i);
            V = r.V; w = r.w; g_ex = r.g_ex; dg_ex = r.dg_ex; g_in = r.g_in; dg_in = r.dg_in; refr = r.refractory;
        } else {
            const NeuronConst32& p = exc ? Q.ne : Q.ni;
            float exc_input = Q.w_e_abs * exc_sum;
            const float inh_input = Q.w_i_abs * inh_sum;
            if (ev > 0) exc_input += (exc ? Q.w_poiss_exc : Q.w_poiss_inh) * static_cast<float>(ev);
            const float I_ext = I_sic + noise_val;
            NeuronRegs32 r{V, w, g_ex, dg_ex + exc_input * p.psc_init_ex, g_in, dg_in + inh_input * p.psc_init_in, refr};
            spiked = neuron_substeps32(p, Q, r, I_ext, V_reset, bb, P.substeps, exc ? Q.ex_decay_e : Q.ex_decay_i,
                                       exc ? Q.in_decay_e : Q.in_decay_i);
            V = r.V; w = r.w; g_ex = r.g_ex; dg_ex = r.dg_ex; g_in = r.g_in; dg_in = r.dg_in; refr = r.refractory;
        }
        if (spiked) {
            R out = R(1);
            if (P.stp_enabled) {
                R x = S.stp_x[j], u = S.stp_u[j], tl = S.stp_t[j];
                if constexpr (sizeof(R) == 8) {
                    out = stp_on_spike(x, u, tl, static_cast<double>(t) * P.dt, P.stp_U, P.stp_tau_rec, P.stp_tau_fac);
                } else {
                    const float t_now = static_cast<float>(static_cast<double>(t) * P.dt);
                    const float elapsed = t_now - tl;
                    const float x_decay = __expf(-elapsed / static_cast<float>(P.stp_tau_rec));
                    const float u_decay =
                        P.stp_tau_fac < 1e-10 ? 0.0f : __expf(-elapsed / static_cast<float>(P.stp_tau_fac));
                    const float U = static_cast<float>(P.stp_U);
                    x = 1.0f + (x - x * u - 1.0f) * x_decay;
                    u = U + u * (1.0f - U) * u_decay;
                    tl = t_now;
                    out = x * u;
                }
                S.stp_x[j] = x;
                S.stp_u[j] = u;
                S.stp_t[j] = tl;
            }
            S.spk_v[std::size_t(t % B.Rv) * D.n_neurons + j] = out;
            wout |= 1u << k;
            ++nspikes;
        }
    }
    S.V[j] = V;
    S.w[j] = w;
    S.g_ex[j] = g_ex;
    S.dg_ex[j] = dg_ex;
    S.g_in[j] = g_in;
    S.dg_in[j] = dg_in;
    D.refr[j] = refr;
    S.I_sic[j] = I_sic;
    D.spk_w[std::size_t(B.wslot) * D.n_neurons + j] = wout;
    if (B.last) {
        D.bcnt[T_EXC][j] = 0;
        D.bcnt[T_INH][j] = 0;
        D.bcnt[T_SIC][j] = 0;
    }
    *emit = B.last && wout != 0u;
    return nspikes;
}

template <int MAXT, class R>
__global__ __launch_bounds__(MAXT) void integrate_kernel(BlkCommon D, BlkState<R> S, RunParams P, Params32 Q,
                                                         std::uint64_t seed, BlkLaunch B) {
    if (blockIdx.x < D.astro_blocks) {
        const index_t a = blockIdx.x * blockDim.x + threadIdx.x;
        const bool emit = a < D.n_astro ? astro_cell<R>(D, S, P, Q, seed, B, a) : false;
        wave_append(emit, a, D.alist + std::size_t(B.wslot) * D.n_astro, &D.lcount[B.wslot * 2 + 1]);
        return;
    }
    const index_t j = (blockIdx.x - D.astro_blocks) * blockDim.x + threadIdx.x;
    bool emit = false;
    unsigned nspikes = 0;
    if (j < D.n_neurons) nspikes = neuron_cell<R>(D, S, P, Q, seed, B, j, &emit);
    wave_append(emit, j, D.nlist + std::size_t(B.wslot) * D.n_neurons, &D.lcount[B.wslot * 2 + 0]);
    // diagnostic spike total: one atomic per wave with spikes
    unsigned long long c = nspikes;
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off);
    if (lane_id() == 0 && c != 0ull) atomicAdd(&D.counters[blockIdx.x & (kSpikeCounters - 1)], c);
}

// ---------------------------------------------------------------------------
// Setup kernels: transpose an in-CSR into (target, slot) lists per source.
// ---------------------------------------------------------------------------
__global__ void count_sources_kernel(const index_t* src, std::size_t n_edges, index_t n_src, index_t* cnt) {
    const std::size_t e = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < n_edges) {
        const index_t s = src[e];
        if (s < n_src) atomicAdd(&cnt[s], 1u);
    }
}

__global__ void fill_out_kernel(const index_t* rs, const index_t* src, index_t n_targets, index_t n_src,
                                index_t* cursor, uint2* out) {
    const index_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_targets) return;
    for (index_t e = rs[t]; e < rs[t + 1]; ++e) {
        const index_t s = src[e];
        if (s < n_src) out[atomicAdd(&cursor[s], 1u)] = make_uint2(t, e);
    }
}


// Block geometry shared by the engine and its footprint.
struct Geometry {
    int d = 10, L = 10, H = 2, Rv = 20;
};

Geometry geometry(const RunParams& rp) {
    if (rp.d_e != rp.d_i || rp.d_e != rp.d_a2n)
        throw Error(AS_ERR_INVALID, "the blocked method needs equal exc/inh/a2n delays");
    Geometry g;
    g.d = rp.d_e;
    g.L = 1;
    for (int l = std::min(g.d, 32); l >= 1; --l)
        if (g.d % l == 0) { g.L = l; break; }
    g.H = g.d / g.L + 1;
    g.Rv = g.H * g.L;
    return g;
}

struct BlkLayout {
    std::size_t in_rs[4], in_src[4], out_rs[4], out_ent[4], bcnt[4], bent[4];
    std::size_t astro[10], ever, neuron[12], refr;
    std::size_t spk_w, spk_v, sic_w, sic_v, nlist, alist, lcount, counters, overflow;
    std::size_t total;
};

index_t out_sources(const Network& net, int T) {
    switch (T) {
        case T_EXC: return net.cfg.n_exc;
        case T_INH: return net.cfg.n_inh;  // local inhibitory ids index neurons 0..n_inh-1
        case T_SIC: return net.cfg.n_astro;
        default: return net.cfg.n_exc;
    }
}

// rsize: bytes of one state value (8 FP64, 4 FP32c).
BlkLayout blk_layout(const Network& net, const RunParams& rp, std::size_t rsize) {
    const Geometry g = geometry(rp);
    Layout L;
    BlkLayout o{};
    const HostCSR* c[4] = {&net.inc_exc, &net.inc_inh, &net.inc_sic, &net.inc_e2a};
    const std::size_t na = net.cfg.n_astro, nn = net.cfg.n_neurons();
    for (int i = 0; i < 4; ++i) {
        const std::size_t ne = std::max<std::size_t>(1, c[i]->src.size());
        o.in_rs[i] = L.add(c[i]->row_start.size() * sizeof(index_t));
        o.in_src[i] = L.add(ne * sizeof(index_t));
        o.out_rs[i] = L.add((std::size_t(out_sources(net, i)) + 1) * sizeof(index_t));
        o.out_ent[i] = L.add(ne * sizeof(uint2));
        const std::size_t nt = c[i]->row_start.size() - 1;
        o.bcnt[i] = L.add(nt * sizeof(unsigned));
        o.bent[i] = L.add(nt * kBucketCap * sizeof(uint2));
    }
    // Ca IP3 h, Ca_tot IP3_0 tau delta, and (FP32c) the three compensation terms
    const int n_astro_arrays = rsize == 4 ? 10 : 7;
    for (int i = 0; i < n_astro_arrays; ++i) o.astro[i] = L.add(na * rsize);
    o.ever = L.add(na);
    for (int i = 0; i < 12; ++i) o.neuron[i] = L.add(nn * rsize);
    o.refr = L.add(nn * sizeof(std::int32_t));
    o.spk_w = L.add(std::size_t(g.H) * nn * sizeof(std::uint32_t));
    o.spk_v = L.add(std::size_t(g.Rv) * nn * rsize);
    o.sic_w = L.add(std::size_t(g.H) * na * sizeof(std::uint32_t));
    o.sic_v = L.add(std::size_t(g.Rv) * na * rsize);
    o.nlist = L.add(std::size_t(g.H) * nn * sizeof(std::uint32_t));
    o.alist = L.add(std::size_t(g.H) * na * sizeof(std::uint32_t));
    o.lcount = L.add(std::size_t(g.H) * 2 * sizeof(std::uint32_t));
    o.counters = L.add(kSpikeCounters * sizeof(unsigned long long));
    o.overflow = L.add(sizeof(unsigned));
    o.total = L.total;
    return o;
}

template <class R>
__global__ void fill_kernel(R* p, R v, std::size_t n) {
    const std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <class R>
class BlockedEngine final : public Engine {
public:
    BlockedEngine(const Network& net, const RunParams& rp, hipStream_t s)
        : P_(rp), Q_(make_params32(rp)), seed_(net.seed) {
        stream_ = s;
        geo_ = geometry(rp);
        n_neurons_ = net.cfg.n_neurons();
        n_astro_ = net.cfg.n_astro;
        n_exc_ = net.cfg.n_exc;
        lay_ = blk_layout(net, rp, sizeof(R));
        if (hipMalloc(&base_, lay_.total) != hipSuccess) {
            (void)hipGetLastError();
            base_ = nullptr;
            throw Error(AS_ERR_NOMEM, "device allocation of " + std::to_string(lay_.total) + " bytes failed");
        }
        char* b = static_cast<char*>(base_);
        auto dptr = [&](std::size_t off) { return static_cast<void*>(b + off); };
        const HostCSR* c[4] = {&net.inc_exc, &net.inc_inh, &net.inc_sic, &net.inc_e2a};
        for (int i = 0; i < 4; ++i) {
            AS_HIP_CHECK(hipMemcpy(dptr(lay_.in_rs[i]), c[i]->row_start.data(),
                                   c[i]->row_start.size() * sizeof(index_t), hipMemcpyHostToDevice));
            if (!c[i]->src.empty())
                AS_HIP_CHECK(hipMemcpy(dptr(lay_.in_src[i]), c[i]->src.data(), c[i]->src.size() * sizeof(index_t),
                                       hipMemcpyHostToDevice));
            D_.in_rs[i] = (const index_t*)dptr(lay_.in_rs[i]);
            D_.in_src[i] = (const index_t*)dptr(lay_.in_src[i]);
            D_.out_rs[i] = (const index_t*)dptr(lay_.out_rs[i]);
            D_.out_ent[i] = (const uint2*)dptr(lay_.out_ent[i]);
            D_.out_nsrc[i] = out_sources(net, i);
            D_.bcnt[i] = (unsigned*)dptr(lay_.bcnt[i]);
            D_.bent[i] = (uint2*)dptr(lay_.bent[i]);
            const index_t nt = c[i]->n_targets();
            AS_HIP_CHECK(hipMemsetAsync(D_.bcnt[i], 0, std::size_t(nt) * sizeof(unsigned), stream_));
            build_out_csr(i, nt, c[i]->n_edges());
        }
        R* A[10] = {};
        for (int i = 0; i < (sizeof(R) == 4 ? 10 : 7); ++i) A[i] = (R*)dptr(lay_.astro[i]);
        S_.Ca = A[0]; S_.IP3 = A[1]; S_.h = A[2];
        S_.Ca_tot = A[3]; S_.IP3_0 = A[4]; S_.tau_IP3 = A[5]; S_.delta_IP3 = A[6];
        S_.Ca_lo = A[7]; S_.IP3_lo = A[8]; S_.h_lo = A[9];
        D_.ever = (std::uint8_t*)dptr(lay_.ever);
        R* N[12];
        for (int i = 0; i < 12; ++i) N[i] = (R*)dptr(lay_.neuron[i]);
        S_.V = N[0]; S_.w = N[1]; S_.g_ex = N[2]; S_.dg_ex = N[3]; S_.g_in = N[4]; S_.dg_in = N[5];
        S_.I_sic = N[6]; S_.stp_x = N[7]; S_.stp_u = N[8]; S_.stp_t = N[9];
        S_.V_reset = N[10]; S_.b = N[11];
        D_.refr = (std::int32_t*)dptr(lay_.refr);
        D_.spk_w = (std::uint32_t*)dptr(lay_.spk_w);
        S_.spk_v = (R*)dptr(lay_.spk_v);
        D_.sic_w = (std::uint32_t*)dptr(lay_.sic_w);
        S_.sic_v = (R*)dptr(lay_.sic_v);
        D_.nlist = (std::uint32_t*)dptr(lay_.nlist);
        D_.alist = (std::uint32_t*)dptr(lay_.alist);
        D_.lcount = (std::uint32_t*)dptr(lay_.lcount);
        D_.counters = (unsigned long long*)dptr(lay_.counters);
        D_.overflow = (unsigned*)dptr(lay_.overflow);
        D_.n_astro = n_astro_;
        D_.n_neurons = n_neurons_;
        D_.n_exc = n_exc_;
        D_.n_inh = net.cfg.n_inh;
        D_.astro_blocks = (n_astro_ + rp.block - 1) / rp.block;
        neuron_blocks_ = (n_neurons_ + rp.block - 1) / rp.block;

        // Parameters and initial state (src/astrosim_fused.cu:981-1018). In
        // FP32c the astrocyte initial values keep their rounding residue in
        // the compensation terms and tau_IP3 is stored as its reciprocal.
        const Biology bio;
        upload(A[3], net.Ca_tot);
        upload(A[4], net.IP3_0);
        if constexpr (sizeof(R) == 4) {
            std::vector<double> inv(net.tau_IP3.size());
            for (std::size_t i = 0; i < inv.size(); ++i) inv[i] = 1.0 / net.tau_IP3[i];
            upload(A[5], inv);
        } else {
            upload(A[5], net.tau_IP3);
        }
        upload(A[6], net.delta_IP3);
        upload(N[10], net.V_reset);
        upload(N[11], net.b);
        fill(A[0], R(bio.Ca_init), n_astro_);
        fill(A[1], R(bio.IP3_init), n_astro_);
        fill(A[2], R(bio.h_init), n_astro_);
        if constexpr (sizeof(R) == 4) {
            fill(A[7], R(bio.Ca_init - double(float(bio.Ca_init))), n_astro_);
            fill(A[8], R(bio.IP3_init - double(float(bio.IP3_init))), n_astro_);
            fill(A[9], R(bio.h_init - double(float(bio.h_init))), n_astro_);
        }
        {
            std::vector<double> V0(n_neurons_);
            for (index_t i = 0; i < n_neurons_; ++i) V0[i] = i < n_exc_ ? bio.e_V_m : bio.i_V_m;
            upload(N[0], V0);
        }
        for (int i = 1; i <= 6; ++i) AS_HIP_CHECK(hipMemsetAsync(N[i], 0, n_neurons_ * sizeof(R), stream_));
        fill(N[7], R(1), n_neurons_);
        fill(N[8], R(rp.stp_U), n_neurons_);
        fill(N[9], R(-1e12), n_neurons_);
        AS_HIP_CHECK(hipMemsetAsync(D_.refr, 0, n_neurons_ * sizeof(std::int32_t), stream_));
        AS_HIP_CHECK(hipMemsetAsync(D_.ever, 0, n_astro_, stream_));
        AS_HIP_CHECK(hipMemsetAsync(D_.spk_w, 0, std::size_t(geo_.H) * n_neurons_ * 4, stream_));
        AS_HIP_CHECK(hipMemsetAsync(D_.sic_w, 0, std::size_t(geo_.H) * n_astro_ * 4, stream_));
        AS_HIP_CHECK(hipMemsetAsync(S_.spk_v, 0, std::size_t(geo_.Rv) * n_neurons_ * sizeof(R), stream_));
        AS_HIP_CHECK(hipMemsetAsync(S_.sic_v, 0, std::size_t(geo_.Rv) * n_astro_ * sizeof(R), stream_));
        AS_HIP_CHECK(hipMemsetAsync(D_.lcount, 0, std::size_t(geo_.H) * 2 * 4, stream_));
        AS_HIP_CHECK(hipMemsetAsync(D_.counters, 0, kSpikeCounters * sizeof(unsigned long long), stream_));
        AS_HIP_CHECK(hipMemsetAsync(D_.overflow, 0, sizeof(unsigned), stream_));
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    ~BlockedEngine() override {
        if (base_) {
            (void)hipStreamSynchronize(stream_);
            (void)hipFree(base_);
        }
    }

    void enqueue_steps(std::int64_t k) override {
        if (k <= 0) return;
        const unsigned grid = D_.astro_blocks + neuron_blocks_;
        while (k > 0) {
            const std::int64_t bidx = step_ / geo_.L;
            BlkLaunch B;
            B.bstart = bidx * geo_.L;
            B.t0 = step_;
            B.n = static_cast<int>(std::min<std::int64_t>(k, B.bstart + geo_.L - step_));
            B.d = geo_.d;
            B.Rv = geo_.Rv;
            B.wslot = static_cast<std::uint32_t>(bidx % geo_.H);
            B.rslot = static_cast<std::uint32_t>(((bidx - geo_.d / geo_.L) % geo_.H + geo_.H) % geo_.H);
            B.first = step_ == B.bstart;
            B.last = step_ + B.n == B.bstart + geo_.L;
            if (B.first) {
                hipLaunchKernelGGL(scatter_kernel, dim3(kScatterBlocks), dim3(256), 0, stream_, D_, B.rslot, B.wslot);
            }
            timing_begin();  // the integrate kernel is the dominant step kernel
            if (P_.block <= 256)
                hipLaunchKernelGGL((integrate_kernel<256, R>), dim3(grid), dim3(P_.block), 0, stream_, D_, S_, P_, Q_,
                                   seed_, B);
            else
                hipLaunchKernelGGL((integrate_kernel<1024, R>), dim3(grid), dim3(P_.block), 0, stream_, D_, S_, P_, Q_,
                                   seed_, B);
            timing_end(1);
            step_ += B.n;
            k -= B.n;
        }
        AS_HIP_CHECK(hipGetLastError());
    }

    std::uint64_t spike_count() override {
        unsigned long long h[kSpikeCounters];
        AS_HIP_CHECK(hipMemcpyAsync(h, D_.counters, sizeof h, hipMemcpyDeviceToHost, stream_));
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
        std::uint64_t t = 0;
        for (int i = 0; i < kSpikeCounters; ++i) t += h[i];
        return t;
    }

    void reset_ever_active() override { AS_HIP_CHECK(hipMemsetAsync(D_.ever, 0, n_astro_, stream_)); }

    void read_state(int field, double* out) override {
        const R* src = nullptr;
        const R* lo = nullptr;
        std::size_t n = n_neurons_;
        switch (field) {
            case 0: src = S_.Ca; lo = S_.Ca_lo; n = n_astro_; break;
            case 1: src = S_.IP3; lo = S_.IP3_lo; n = n_astro_; break;
            case 2: src = S_.h; lo = S_.h_lo; n = n_astro_; break;
            case 3: src = S_.V; break;
            case 4: src = S_.w; break;
            case 5: src = S_.g_ex; break;
            case 6: src = S_.dg_ex; break;
            case 7: src = S_.g_in; break;
            case 8: src = S_.dg_in; break;
            case 9: src = S_.I_sic; break;
            case 10: src = S_.stp_x; break;
            case 11: src = S_.stp_u; break;
            case 12: src = S_.stp_t; break;
            default: throw Error(AS_ERR_INVALID, "unknown state field " + std::to_string(field));
        }
        std::vector<R> h(n), l;
        AS_HIP_CHECK(hipMemcpyAsync(h.data(), src, n * sizeof(R), hipMemcpyDeviceToHost, stream_));
        if (sizeof(R) == 4 && lo) {
            l.resize(n);
            AS_HIP_CHECK(hipMemcpyAsync(l.data(), lo, n * sizeof(R), hipMemcpyDeviceToHost, stream_));
        }
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
        for (std::size_t i = 0; i < n; ++i) out[i] = double(h[i]) + (l.empty() ? 0.0 : double(l[i]));
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
        // Block words of the last H blocks are held.
        const std::int64_t cur_block = step_ > 0 ? (step_ - 1) / geo_.L : 0;
        if (n < 0 || first < 0 || first + n > step_ || (n > 0 && first / geo_.L <= cur_block - geo_.H))
            throw Error(AS_ERR_RANGE, "raster steps not held (the blocked method keeps the last H blocks)");
        std::vector<std::uint32_t> words(n_neurons_);
        std::int64_t have_block = -1;
        for (int i = 0; i < n; ++i) {
            const std::int64_t s = first + i;
            const std::int64_t bl = s / geo_.L;
            if (bl != have_block) {
                AS_HIP_CHECK(hipMemcpyAsync(words.data(), D_.spk_w + std::size_t(bl % geo_.H) * n_neurons_,
                                            n_neurons_ * 4, hipMemcpyDeviceToHost, stream_));
                AS_HIP_CHECK(hipStreamSynchronize(stream_));
                have_block = bl;
            }
            const unsigned bit = static_cast<unsigned>(s - bl * geo_.L);
            for (index_t j = 0; j < n_neurons_; ++j) out[std::size_t(i) * n_neurons_ + j] = (words[j] >> bit) & 1u;
        }
    }

    std::uint64_t device_bytes() const override { return lay_.total; }

private:
    void upload(const R* dst, const std::vector<double>& v) {
        if (v.empty()) return;
        if constexpr (sizeof(R) == 8) {
            AS_HIP_CHECK(hipMemcpy(const_cast<R*>(dst), v.data(), v.size() * sizeof(R), hipMemcpyHostToDevice));
        } else {
            std::vector<R> f(v.begin(), v.end());
            AS_HIP_CHECK(hipMemcpy(const_cast<R*>(dst), f.data(), f.size() * sizeof(R), hipMemcpyHostToDevice));
        }
    }

    void fill(R* p, R v, std::size_t n) {
        if (n == 0) return;
        hipLaunchKernelGGL(fill_kernel<R>, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream_, p, v, n);
        AS_HIP_CHECK(hipGetLastError());
    }

    // Transposes in-CSR type T into per-source (target, in-CSR slot) lists:
    // count per source, exclusive scan, then scatter with atomic cursors. The
    // order inside a source's list is irrelevant (targets order by slot).
    void build_out_csr(int T, index_t n_targets, std::size_t n_edges) {
        const index_t n_src = D_.out_nsrc[T];
        const std::size_t nrs = std::size_t(n_src) + 1;
        index_t* out_rs = const_cast<index_t*>(D_.out_rs[T]);
        AS_HIP_CHECK(hipMemsetAsync(out_rs, 0, nrs * sizeof(index_t), stream_));
        if (n_edges > 0)
            hipLaunchKernelGGL(count_sources_kernel, dim3(unsigned((n_edges + 255) / 256)), dim3(256), 0, stream_,
                               D_.in_src[T], n_edges, n_src, out_rs);
        AS_HIP_CHECK(hipGetLastError());
        std::size_t tmp_bytes = 0;
        AS_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, out_rs, out_rs, int(nrs), stream_));
        void* tmp = nullptr;
        index_t* cursor = nullptr;
        AS_HIP_CHECK(hipMalloc(&tmp, std::max<std::size_t>(tmp_bytes, 1)));
        AS_HIP_CHECK(hipMalloc(&cursor, nrs * sizeof(index_t)));
        AS_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, out_rs, out_rs, int(nrs), stream_));
        AS_HIP_CHECK(hipMemcpyAsync(cursor, out_rs, nrs * sizeof(index_t), hipMemcpyDeviceToDevice, stream_));
        if (n_targets > 0)
            hipLaunchKernelGGL(fill_out_kernel, dim3((n_targets + 255) / 256), dim3(256), 0, stream_, D_.in_rs[T],
                               D_.in_src[T], n_targets, n_src, cursor, const_cast<uint2*>(D_.out_ent[T]));
        AS_HIP_CHECK(hipGetLastError());
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
        AS_HIP_CHECK(hipFree(tmp));
        AS_HIP_CHECK(hipFree(cursor));
    }

    static constexpr unsigned kScatterBlocks = 1024;
    RunParams P_;
    Params32 Q_;
    std::uint64_t seed_;
    Geometry geo_;
    BlkCommon D_{};
    BlkState<R> S_{};
    BlkLayout lay_{};
    void* base_ = nullptr;
    std::uint32_t neuron_blocks_ = 0;
};

}  // namespace

std::uint64_t blocked_footprint(const Network& net, const RunParams& rp) {
    return blk_layout(net, rp, rp.precision == AS_FP64 ? 8 : 4).total;
}

std::unique_ptr<Engine> make_blocked_engine(const Network& net, const RunParams& rp, hipStream_t s) {
    if (rp.precision == AS_FP64) return std::unique_ptr<Engine>(new BlockedEngine<double>(net, rp, s));
    return std::unique_ptr<Engine>(new BlockedEngine<float>(net, rp, s));
}

}  // namespace astrosim
