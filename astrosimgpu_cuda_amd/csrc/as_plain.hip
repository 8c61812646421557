// as_plain.hip -- the plain FP64 method: every in-edge visited every step.
//
// This is the reference's data flow (src/astrosim_fused.cu:374-587): one lane
// per cell for the cell arithmetic, each cell summing its whole in-CSR rows
// at the delayed step, writing only its own state and its own delay-ring
// entry. The astrocyte and neuron phases of a step read only ring slots
// written in earlier steps (every delay is >= 1 step and < R), so they need
// no ordering between them inside a step.
//
// What differs from the reference is how the delayed outputs are held and
// how a row is walked, because every step streams the whole CSR (about 5.3
// GB of source ids at N_total = 10M) and that stream is the cost:
//   - Delay rings are bit-packed: one 64-bit word of spike flags per wave of
//     64 neurons per step (a wave ballot), plus the FP64 release value x*u
//     written only by the neurons that spiked. Astrocytes likewise keep one
//     word of "SIC factor > 0" flags per 64 cells and the factor where it is
//     nonzero. A zero ring entry contributes an exact +0.0 to the reference's
//     sums (every term is >= +0.0 and the sums start at +0.0), so summing only
//     the flagged terms, in row order, gives the reference's sums bit for bit.
//   - Each workgroup stages the read slot's flags in LDS: the flag words
//     themselves when they fit the LDS budget (exact), otherwise a summary
//     holding one bit per 2^gsh cells (gsh >= 6), set by an atomic OR from
//     every wave that produced a nonzero flag word; a set summary bit is then
//     confirmed against the flag word in global memory (L2-resident: 1.25 MB
//     per step at 10M neurons).
//   - A wave gathers the in-rows of its 64 targets together: their source
//     ids are one contiguous range of the CSR, which the wave streams in
//     coalesced 16-byte-per-lane loads (4 KB per wave per round, the next
//     round in flight while the current one is tested against the LDS bits
//     and confirmed). The rare flagged terms are then added in position
//     order, each to the lane whose row holds it (found by a ballot over the
//     wave's row starts), which is row order.
// Two launch shapes:
//   - FUSED (small networks, launch-bound): one kernel per step; each
//     workgroup stages the bits once and loops over tiles of cells, gathering
//     and integrating each wave.
//   - SPLIT (from 2^16 neurons): a gather kernel (80 registers, 6 waves per
//     SIMD keep the id stream in flight) leaves each target's input sums
//     where nonzero plus one word per wave saying which; a cell kernel (one
//     tile per workgroup) runs the arithmetic. The fused kernel's cell
//     arithmetic holds 128 registers, which caps the waves streaming ids at 4
//     per SIMD.
#include "as_engine.h"
#include "as_model.h"

namespace astrosim {
namespace {

constexpr int kSpikeCounters = 64;
constexpr int kLoadsInFlight = 4;  // 16-byte id loads per lane per gather round

struct PlainDev {
    const index_t *exc_rs, *exc_src, *inh_rs, *inh_src, *sic_rs, *sic_src, *e2a_rs, *e2a_src;
    double *Ca, *IP3, *h;
    const double *Ca_tot, *IP3_0, *tau_IP3, *delta_IP3;
    std::uint8_t* ever;
    double *V, *w, *g_ex, *dg_ex, *g_in, *dg_in, *I_sic;
    std::int32_t* refr;
    std::uint32_t* spk_n;  // spikes per neuron since creation
    const double *V_reset, *b;
    double *stp_x, *stp_u, *stp_t;
    std::uint64_t *flag_n, *flag_s;  // [R][wn], [R][wa] flag words per 64 cells
    double *val_n, *val_s;           // [R][n_neurons], [R][n_astro], valid where flagged
    std::uint32_t *sum_n, *sum_s;    // [R+1][sn], [R+1][sa] summary bits (summary mode)
    unsigned long long* counters;    // kSpikeCounters partial spike totals
    double *in_exc, *in_inh, *in_sic, *in_e2a;  // split mode: input sums where nonzero
    std::uint64_t *mask_exc, *mask_inh, *mask_sic, *mask_e2a;  // split mode: which lanes have one
    index_t n_astro, n_neurons, n_exc;
    std::uint32_t wn, wa;            // flag words per slot
    std::uint32_t sn, sa;            // LDS words staged per population
    int gsh_n, gsh_s;                // cells per staged bit = 2^gsh (0: exact flags)
    std::uint32_t astro_waves, neuron_waves;  // 64-cell waves of each population
};

struct PlainStep {
    std::int64_t step;
    std::uint64_t noise_idx;
    int sic_refresh;
    std::uint32_t rd_n, rd_s, wr;       // flag / value ring slots (R deep)
    std::uint32_t srd_n, srd_s, swr, sclr;  // summary ring slots (R + 1 deep)
};

__device__ __forceinline__ unsigned lane_id() { return threadIdx.x & 63u; }

__device__ __forceinline__ std::uint32_t uniform(std::uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ std::uint32_t lane_u32(std::uint32_t x, unsigned l) {
    return static_cast<std::uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), static_cast<int>(l)));
}

// A slot of 64-bit flag words read as 32-bit words (little-endian: cell i
// is bit i & 31 of word i >> 5).
__device__ __forceinline__ const std::uint32_t* words32(const std::uint64_t* p) {
    return reinterpret_cast<const std::uint32_t*>(p);
}

// Whether cell `id` may be flagged, from the LDS bits: exact when gsh == 0.
__device__ __forceinline__ bool lds_bit(const std::uint32_t* bits, int gsh, index_t id) {
    const index_t s = id >> gsh;
    return (bits[s >> 5] >> (s & 31u)) & 1u;
}

// Sum of the flagged delayed values over the in-rows of a wave's targets.
// The rows are [rs_lane, next row) of the wave's valid lanes and together
// form the contiguous CSR range [lo, hi) (wave-uniform); lanes past the last
// valid target hold rs_lane = hi. Each flagged term is added to its row's
// lane in position order (the reference's row order, src:384-390, 450-463,
// 482-490, with its exact +0.0 terms left out).
//
// The range is streamed in rounds of U 16-byte loads per lane (U KB per
// wave), software-pipelined: while a round's candidates are confirmed
// against the flag words (one L2 round trip in summary mode), the next
// round's ids are already in flight.
template <int U>
__device__ __forceinline__ void load_round(uint4 (&v)[U], const index_t* __restrict__ src, index_t base, index_t hi) {
    const unsigned lane = lane_id();
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const index_t p = base + 256u * u + 4u * lane;
        v[u] = p < hi ? *reinterpret_cast<const uint4*>(src + p) : make_uint4(0u, 0u, 0u, 0u);
    }
}

template <int U>
__device__ __forceinline__ double wave_gather(const index_t* __restrict__ src, index_t lo, index_t hi, index_t rs_lane,
                              const std::uint32_t* bits, int gsh, const std::uint32_t* __restrict__ flags,
                              const double* __restrict__ vals) {
    static_assert(U >= 1 && U <= 8, "candidate mask holds 4*U bits");
    double acc = 0.0;
    const unsigned lane = lane_id();
    const bool exact = gsh == 0;
    index_t base = lo & ~3u;
    if (base >= hi) return acc;
    uint4 cur[U];
    load_round<U>(cur, src, base, hi);
    for (;;) {
        const index_t next = base + 256u * U;
        // candidate bits: bit 4*u + c for position base + 256u + 4*lane + c
        std::uint32_t cand = 0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const index_t p = base + 256u * u + 4u * lane;
            const index_t id4[4] = {cur[u].x, cur[u].y, cur[u].z, cur[u].w};
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const index_t pos = p + c;
                const bool in = pos >= lo && pos < hi;
                if (in && lds_bit(bits, gsh, id4[c])) cand |= 1u << (4 * u + c);
            }
        }
        // summary mode: the flag words of the candidates, loaded before the
        // next round's ids so that waiting for them does not wait for those
        std::uint32_t fw[4 * U];
        if (!exact) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const index_t id4[4] = {cur[u].x, cur[u].y, cur[u].z, cur[u].w};
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int k = 4 * u + c;
                    fw[k] = ((cand >> k) & 1u) ? flags[id4[c] >> 5] : 0u;
                }
            }
        }
        uint4 nxt[U];
        const bool more = next < hi;
        if (more) load_round<U>(nxt, src, next, hi);
        if (!exact) {
            std::uint32_t keep = 0;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const index_t id4[4] = {cur[u].x, cur[u].y, cur[u].z, cur[u].w};
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int k = 4 * u + c;
                    if ((fw[k] >> (id4[c] & 31u)) & 1u) keep |= 1u << k;
                }
            }
            cand &= keep;
        }
        if (__ballot(cand != 0u) != 0ull) {
            // flagged terms, in position order: load by load, lane by lane, component by component
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const index_t id4[4] = {cur[u].x, cur[u].y, cur[u].z, cur[u].w};
                unsigned long long mu = __ballot(((cand >> (4 * u)) & 15u) != 0u);
                while (mu != 0ull) {
                    const unsigned l = __ffsll(static_cast<long long>(mu)) - 1;
                    mu &= mu - 1ull;
                    const std::uint32_t cl = lane_u32(cand, l) >> (4 * u);
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        if ((cl >> c) & 1u) {
                            const index_t id = lane_u32(id4[c], l);
                            const index_t pos = base + 256u * u + 4u * l + c;
                            const double val = vals[id];
                            const unsigned t = __popcll(__ballot(rs_lane <= pos)) - 1u;
                            if (lane == t) acc += val;
                        }
                    }
                }
            }
        }
        if (!more) break;
        base = next;
#pragma unroll
        for (int u = 0; u < U; ++u) cur[u] = nxt[u];
    }
    return acc;
}

// Row bounds of a wave's targets j0 + lane in one CSR: per lane its row
// start (hi past the last valid target), and the wave's range [lo, hi).
struct WaveRows {
    index_t rs, lo, hi;
};

__device__ __forceinline__ WaveRows wave_rows(const index_t* __restrict__ row_start, index_t j0, index_t nv) {
    WaveRows r;
    const index_t lane = lane_id();
    r.lo = uniform(row_start[j0]);
    r.hi = uniform(row_start[j0 + nv]);
    r.rs = row_start[j0 + (lane < nv ? lane : nv)];
    return r;
}

// Publishes a wave's 64 flags for its step: the flag word, the summary bit
// (summary mode) and the spike/SIC value of every flagged lane.
__device__ __forceinline__ void wave_publish(bool flag, double value, index_t j, index_t j0, std::uint64_t* flag_slot,
                                             double* val_slot, std::uint32_t* sum_slot, int gsh) {
    const unsigned long long m = __ballot(flag);
    if (flag) val_slot[j] = value;
    if (lane_id() == 0) {
        flag_slot[j0 >> 6] = m;
        if (gsh != 0 && m != 0ull) {
            const index_t s = j0 >> gsh;
            atomicOr(&sum_slot[s >> 5], 1u << (s & 31u));
        }
    }
}

// Workgroup tiles of `waves` waves: astrocyte tiles first, then neurons.
struct Tiles {
    std::uint32_t astro, total;
};

__host__ __device__ inline Tiles make_tiles(std::uint32_t astro_waves, std::uint32_t neuron_waves, std::uint32_t waves) {
    const std::uint32_t a = (astro_waves + waves - 1) / waves;
    return Tiles{a, a + (neuron_waves + waves - 1) / waves};
}

// Clears the summary slot of the next step (one workgroup's threads).
__device__ __forceinline__ void clear_next_summary(const PlainDev& D, const PlainStep& S) {
    if (D.gsh_n != 0)
        for (std::uint32_t i = threadIdx.x; i < D.sn; i += blockDim.x) D.sum_n[std::size_t(S.sclr) * D.sn + i] = 0u;
    if (D.gsh_s != 0)
        for (std::uint32_t i = threadIdx.x; i < D.sa; i += blockDim.x) D.sum_s[std::size_t(S.sclr) * D.sa + i] = 0u;
}


// The n2a input of a wave of astrocytes a0 + lane (src/astrosim_fused.cu:384-390).
template <int U>
__device__ __forceinline__ double astro_sum(const PlainDev& D, const PlainStep& S, const std::uint32_t* lds_n,
                                            index_t a0, index_t nv) {
    const WaveRows r = wave_rows(D.e2a_rs, a0, nv);
    return wave_gather<U>(D.e2a_src, r.lo, r.hi, r.rs, lds_n, D.gsh_n, words32(D.flag_n + std::size_t(S.rd_n) * D.wn),
                          D.val_n + std::size_t(S.rd_n) * D.n_neurons);
}

// The excitatory, inhibitory and SIC inputs of a wave of neurons j0 + lane
// (src/astrosim_fused.cu:450-463, 482-490).
template <int U>
__device__ __forceinline__ void neuron_sums(const PlainDev& D, const PlainStep& S, const std::uint32_t* lds_n,
                                            const std::uint32_t* lds_s, index_t j0, index_t nv, double& exc_sum,
                                            double& inh_sum, double& sic_sum) {
    const std::uint32_t* fn = words32(D.flag_n + std::size_t(S.rd_n) * D.wn);
    const double* vn = D.val_n + std::size_t(S.rd_n) * D.n_neurons;
    WaveRows r = wave_rows(D.exc_rs, j0, nv);
    exc_sum = wave_gather<U>(D.exc_src, r.lo, r.hi, r.rs, lds_n, D.gsh_n, fn, vn);
    // Inhibitory source ids are local indices and index the neuron ring as
    // they are (src/astrosim_fused.cu:459-462).
    r = wave_rows(D.inh_rs, j0, nv);
    inh_sum = wave_gather<U>(D.inh_src, r.lo, r.hi, r.rs, lds_n, D.gsh_n, fn, vn);
    sic_sum = 0.0;
    if (S.sic_refresh) {
        r = wave_rows(D.sic_rs, j0, nv);
        sic_sum = wave_gather<U>(D.sic_src, r.lo, r.hi, r.rs, lds_s, D.gsh_s,
                                 words32(D.flag_s + std::size_t(S.rd_s) * D.wa), D.val_s + std::size_t(S.rd_s) * D.n_astro);
    }
}

// The arithmetic of astrocyte a once its n2a input is known
// (src/astrosim_fused.cu:392-423). Returns its SIC factor.
__device__ __forceinline__ double astro_cell(const PlainDev& D, const RunParams& P, std::uint64_t seed,
                                             const PlainStep& S, index_t a, double gathered) {
    double ip3_input = P.w_n2a * gathered;
    const std::uint64_t stream = std::uint64_t(S.step) * kStepStride + a;
    if (P.poiss_astro) {
        const int ev = rng_poisson(seed ^ kKeyAstroPoisson, stream, P.lam_astro, P.lim_astro);
        if (ev > 0) ip3_input += P.w_poiss_astro * static_cast<double>(ev);
    }
    const double noise = P.noise_astro > 0.0 ? P.noise_astro * rng_normal(seed ^ kKeyAstroNoise, stream) : 0.0;
    double ca = D.Ca[a], ip3 = D.IP3[a], hh = D.h[a];
    astro_advance(P.ac, D.Ca_tot[a], D.IP3_0[a], make_fixed_divisor(D.tau_IP3[a]), D.delta_IP3[a], ip3_input, noise,
                  P.h_step, P.h_sixth, P.substeps, ca, ip3, hh);
    D.Ca[a] = ca;
    D.IP3[a] = ip3;
    D.h[a] = hh;
    const double factor = sic_factor(P.ac, ca);
    if (factor > 0.0) D.ever[a] = 1;
    return factor;
}

// The arithmetic of one neuron of the excitatory (EXC) or inhibitory
// population once its delayed input sums are known (src/astrosim_fused.cu:
// 465-587). Returns whether it spiked; `out` receives its release value.
template <bool EXC>
__device__ __forceinline__ bool neuron_cell(const PlainDev& D, const RunParams& P, std::uint64_t seed,
                                            const PlainStep& S, index_t j, double exc_sum, double inh_sum,
                                            double sic_sum, double& out) {
    const NeuronConst& p = EXC ? P.ne : P.ni;
    double exc_input = P.w_e * exc_sum;
    const double inh_input = (-P.w_i) * inh_sum;
    if (EXC ? P.poiss_exc : P.poiss_inh) {
        const int ev = rng_poisson(seed ^ kKeyNeuronPoisson, std::uint64_t(S.step) * kStepStride + j,
                                   EXC ? P.lam_exc : P.lam_inh, EXC ? P.lim_exc : P.lim_inh);
        if (ev > 0) exc_input += (EXC ? P.w_poiss_exc : P.w_poiss_inh) * static_cast<double>(ev);
    }
    double I_sic;
    if (S.sic_refresh) {
        I_sic = P.w_a2n * sic_sum;
        D.I_sic[j] = I_sic;
    } else {
        I_sic = D.I_sic[j];
    }
    double I_ext = I_sic;
    const double nstd = EXC ? P.noise_exc : P.noise_inh;
    if (nstd > 0.0) {
        I_ext += nstd * rng_normal(seed ^ kKeyNeuronNoise, S.noise_idx * kNoiseStride + std::uint64_t(j));
    }
    NeuronRegs g;
    g.V = D.V[j];
    g.w = D.w[j];
    g.g_ex = D.g_ex[j];
    g.dg_ex = D.dg_ex[j] + exc_input * p.psc_init_ex;
    g.g_in = D.g_in[j];
    g.dg_in = D.dg_in[j] + inh_input * p.psc_init_in;
    g.refractory = D.refr[j];
    const int refr0 = g.refractory;
    // V_reset and b are read only where the step uses them (refractory
    // clamp or spike), as in the reference's branches (src:522-552).
    const bool spiked = neuron_substeps_lazy(
        p, g, I_ext, [&]() { return ResetPair{D.V_reset[j], D.b[j]}; }, P.h_step, P.h_sixth, P.substeps,
        EXC ? P.ex_decay_e : P.ex_decay_i, EXC ? P.in_decay_e : P.in_decay_i);
    D.V[j] = g.V;
    D.w[j] = g.w;
    D.g_ex[j] = g.g_ex;
    D.dg_ex[j] = g.dg_ex;
    D.g_in[j] = g.g_in;
    D.dg_in[j] = g.dg_in;
    if (g.refractory != refr0) D.refr[j] = g.refractory;
    if (spiked) {
        D.spk_n[j] += 1u;
        if (P.stp_enabled) {
            double x = D.stp_x[j], u = D.stp_u[j], tl = D.stp_t[j];
            out = stp_on_spike(x, u, tl, static_cast<double>(S.step) * P.dt, P.stp_U, P.stp_tau_rec, P.stp_tau_fac);
            D.stp_x[j] = x;
            D.stp_u[j] = u;
            D.stp_t[j] = tl;
        } else {
            out = 1.0;
        }
    }
    return spiked;
}

// Cell arithmetic and publication of a wave of astrocytes a0 + lane.
__device__ __forceinline__ void astro_wave_cells(const PlainDev& D, const RunParams& P, std::uint64_t seed,
                                                 const PlainStep& S, index_t a0, index_t nv, double gathered) {
    const index_t a = a0 + lane_id();
    double factor = 0.0;
    if (lane_id() < nv) factor = astro_cell(D, P, seed, S, a, gathered);
    wave_publish(factor > 0.0, factor, a, a0, D.flag_s + std::size_t(S.wr) * D.wa,
                 D.val_s + std::size_t(S.wr) * D.n_astro, D.sum_s + std::size_t(S.swr) * D.sa, D.gsh_s);
}

// Cell arithmetic and publication of a wave of neurons j0 + lane.
// Population constants are compile-time selected so that they stay uniform;
// only the one wave straddling n_exc runs both instantiations.
__device__ __forceinline__ void neuron_wave_cells(const PlainDev& D, const RunParams& P, std::uint64_t seed,
                                                  const PlainStep& S, index_t j0, index_t nv, double exc_sum,
                                                  double inh_sum, double sic_sum) {
    const index_t j = j0 + lane_id();
    const bool valid = lane_id() < nv;
    bool spiked = false;
    double out = 0.0;
    const index_t ne = D.n_exc;
    if (j0 + nv <= ne) {
        if (valid) spiked = neuron_cell<true>(D, P, seed, S, j, exc_sum, inh_sum, sic_sum, out);
    } else if (j0 >= ne) {
        if (valid) spiked = neuron_cell<false>(D, P, seed, S, j, exc_sum, inh_sum, sic_sum, out);
    } else {
        if (valid && j < ne) spiked = neuron_cell<true>(D, P, seed, S, j, exc_sum, inh_sum, sic_sum, out);
        if (valid && j >= ne) spiked = neuron_cell<false>(D, P, seed, S, j, exc_sum, inh_sum, sic_sum, out);
    }
    wave_publish(spiked, out, j, j0, D.flag_n + std::size_t(S.wr) * D.wn, D.val_n + std::size_t(S.wr) * D.n_neurons,
                 D.sum_n + std::size_t(S.swr) * D.sn, D.gsh_n);
    // Diagnostic spike count: one atomic per wave that saw a spike, spread
    // over kSpikeCounters addresses (reference: one atomicAdd per spike, src:584).
    const unsigned long long m = __ballot(spiked);
    if (lane_id() == 0 && m != 0ull)
        atomicAdd(&D.counters[(j0 >> 6) & (kSpikeCounters - 1)], (unsigned long long)__popcll(m));
}

// Stages the staged bits of the read slots into LDS; the last workgroup
// also clears the summary slot of the next step, which nobody reads or
// writes in this step (the summary ring is one slot deeper than the delay).
__device__ __forceinline__ void stage_bits(const PlainDev& D, const PlainStep& S, std::uint32_t* lds_n,
                                           std::uint32_t* lds_s, bool clear_next) {
    const std::uint32_t* gn = D.gsh_n == 0 ? words32(D.flag_n + std::size_t(S.rd_n) * D.wn)
                                           : D.sum_n + std::size_t(S.srd_n) * D.sn;
    for (std::uint32_t i = threadIdx.x; i < D.sn; i += blockDim.x) lds_n[i] = gn[i];
    if (S.sic_refresh) {
        const std::uint32_t* gs = D.gsh_s == 0 ? words32(D.flag_s + std::size_t(S.rd_s) * D.wa)
                                               : D.sum_s + std::size_t(S.srd_s) * D.sa;
        for (std::uint32_t i = threadIdx.x; i < D.sa; i += blockDim.x) lds_s[i] = gs[i];
    }
    if (clear_next && blockIdx.x == gridDim.x - 1) clear_next_summary(D, S);
    __syncthreads();
}

// FUSED: one launch per step; each workgroup stages the bits once and loops
// over tiles (astrocyte tiles first), gathering and integrating each wave.
template <int MAXT, int U, int MINW>
__global__ __launch_bounds__(MAXT, MINW) void plain_step_kernel(PlainDev D, RunParams P, std::uint64_t seed, PlainStep S) {
    extern __shared__ std::uint32_t lds[];  // [sn] neuron bits of slot rd_n, then [sa] astrocyte bits of rd_s
    std::uint32_t* lds_n = lds;
    std::uint32_t* lds_s = lds + D.sn;
    stage_bits(D, S, lds_n, lds_s, true);
    const std::uint32_t waves = blockDim.x >> 6;
    const std::uint32_t wave = uniform(threadIdx.x >> 6);
    const Tiles T = make_tiles(D.astro_waves, D.neuron_waves, waves);
    for (std::uint32_t tile = blockIdx.x; tile < T.total; tile += gridDim.x) {
        if (tile < T.astro) {
            const index_t a0 = (tile * waves + wave) * 64u;
            if (a0 < D.n_astro) {
                const index_t nv = uniform(D.n_astro - a0 < 64u ? D.n_astro - a0 : 64u);
                astro_wave_cells(D, P, seed, S, a0, nv, astro_sum<U>(D, S, lds_n, a0, nv));
            }
        } else {
            const index_t j0 = ((tile - T.astro) * waves + wave) * 64u;
            if (j0 < D.n_neurons) {
                const index_t nv = uniform(D.n_neurons - j0 < 64u ? D.n_neurons - j0 : 64u);
                double e, i, c;
                neuron_sums<U>(D, S, lds_n, lds_s, j0, nv, e, i, c);
                neuron_wave_cells(D, P, seed, S, j0, nv, e, i, c);
            }
        }
    }
}

// Writes a wave's input sums where nonzero, and the word of lanes that have one.
__device__ __forceinline__ void put_sum(double v, index_t j, index_t w, double* sums, std::uint64_t* masks) {
    const unsigned long long m = __ballot(v != 0.0);
    if (v != 0.0) sums[j] = v;
    if (lane_id() == 0) masks[w] = m;
}

__device__ __forceinline__ double get_sum(index_t j, index_t w, const double* sums, const std::uint64_t* masks) {
    const std::uint64_t m = masks[w];
    return ((m >> lane_id()) & 1ull) ? sums[j] : 0.0;
}

// SPLIT, first kernel: the gathers alone (few registers, so many waves keep
// the CSR stream in flight). Every target's input sums go to `in` where
// nonzero, flagged per wave in `in_mask`.
template <int MAXT, int U>
__global__ __launch_bounds__(MAXT) void plain_gather_kernel(PlainDev D, PlainStep S) {
    extern __shared__ std::uint32_t lds[];
    std::uint32_t* lds_n = lds;
    std::uint32_t* lds_s = lds + D.sn;
    stage_bits(D, S, lds_n, lds_s, false);
    const std::uint32_t waves = blockDim.x >> 6;
    const std::uint32_t wave = uniform(threadIdx.x >> 6);
    const Tiles T = make_tiles(D.astro_waves, D.neuron_waves, waves);
    for (std::uint32_t tile = blockIdx.x; tile < T.total; tile += gridDim.x) {
        if (tile < T.astro) {
            const index_t a0 = (tile * waves + wave) * 64u;
            if (a0 < D.n_astro) {
                const index_t nv = uniform(D.n_astro - a0 < 64u ? D.n_astro - a0 : 64u);
                const double g = astro_sum<U>(D, S, lds_n, a0, nv);
                put_sum(g, a0 + lane_id(), a0 >> 6, D.in_e2a, D.mask_e2a);
            }
        } else {
            const index_t j0 = ((tile - T.astro) * waves + wave) * 64u;
            if (j0 < D.n_neurons) {
                const index_t nv = uniform(D.n_neurons - j0 < 64u ? D.n_neurons - j0 : 64u);
                double e, i, c;
                neuron_sums<U>(D, S, lds_n, lds_s, j0, nv, e, i, c);
                const index_t j = j0 + lane_id();
                put_sum(e, j, j0 >> 6, D.in_exc, D.mask_exc);
                put_sum(i, j, j0 >> 6, D.in_inh, D.mask_inh);
                put_sum(c, j, j0 >> 6, D.in_sic, D.mask_sic);
            }
        }
    }
}

// SPLIT, second kernel: the cell arithmetic of one tile per workgroup,
// reading the sums the gather kernel left.
template <int MAXT>
__global__ __launch_bounds__(MAXT) void plain_cell_kernel(PlainDev D, RunParams P, std::uint64_t seed, PlainStep S) {
    if (blockIdx.x == gridDim.x - 1) clear_next_summary(D, S);
    const std::uint32_t tile = blockIdx.x;
    const std::uint32_t waves = blockDim.x >> 6;
    const std::uint32_t wave = uniform(threadIdx.x >> 6);
    const Tiles T = make_tiles(D.astro_waves, D.neuron_waves, waves);
    if (tile < T.astro) {
        const index_t a0 = (tile * waves + wave) * 64u;
        if (a0 >= D.n_astro) return;
        const index_t nv = uniform(D.n_astro - a0 < 64u ? D.n_astro - a0 : 64u);
        astro_wave_cells(D, P, seed, S, a0, nv, get_sum(a0 + lane_id(), a0 >> 6, D.in_e2a, D.mask_e2a));
    } else {
        const index_t j0 = ((tile - T.astro) * waves + wave) * 64u;
        if (j0 >= D.n_neurons) return;
        const index_t nv = uniform(D.n_neurons - j0 < 64u ? D.n_neurons - j0 : 64u);
        const index_t j = j0 + lane_id();
        neuron_wave_cells(D, P, seed, S, j0, nv, get_sum(j, j0 >> 6, D.in_exc, D.mask_exc),
                          get_sum(j, j0 >> 6, D.in_inh, D.mask_inh),
                          S.sic_refresh ? get_sum(j, j0 >> 6, D.in_sic, D.mask_sic) : 0.0);
    }
}

__global__ void fill_kernel(double* p, double v, std::size_t n) {
    const std::size_t i = std::size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

struct PlainLayout {
    std::size_t csr_rs[4], csr_src[4];
    std::size_t astro[7], ever;
    std::size_t neuron[12], refr, spk_n;
    std::size_t flag_n, flag_s, val_n, val_s, sum_n, sum_s, counters;
    std::size_t in[4], mask[4];  // split mode: exc, inh, sic, e2a
    bool split;
    std::uint32_t wn, wa, sn, sa;
    int gsh_n, gsh_s;
    std::size_t total;
};

std::uint32_t ceil_div(std::uint64_t a, std::uint64_t b) { return static_cast<std::uint32_t>((a + b - 1) / b); }

// LDS bytes a workgroup may spend on staged flags: enough workgroups of
// `block` threads stay resident for `waves_per_cu` waves per CU (160 KB of
// LDS per CU).
std::size_t lds_budget(int block, int waves_per_cu) {
    const std::size_t b = std::size_t(160 * 1024) * std::size_t(block / 64) / std::size_t(waves_per_cu);
    return std::min<std::size_t>(std::max<std::size_t>(b, 4 * 1024), 64 * 1024);
}

// From this many neurons on, a step is split into a gather kernel and a
// cell kernel (below it one fused launch per step costs less).
constexpr std::uint64_t kSplitMinNeurons = 1u << 16;
// Workgroup size of the split gather kernel (results do not depend on it;
// the cell kernel and the fused kernel use the run's --block).
constexpr int kGatherThreads = 256;

PlainLayout plain_layout(const Network& net, const RunParams& rp) {
    Layout L;
    PlainLayout o{};
    const HostCSR* c[4] = {&net.inc_exc, &net.inc_inh, &net.inc_sic, &net.inc_e2a};
    for (int i = 0; i < 4; ++i) {
        const std::size_t rows = net.on_device() ? std::size_t(net.dev[i].n_targets) + 1 : c[i]->row_start.size();
        o.csr_rs[i] = L.add(rows * sizeof(index_t));
        // 16 bytes of padding: a 16-byte id load may start at the last id
        o.csr_src[i] = L.add(std::max<std::size_t>(1, net.edges(i)) * sizeof(index_t) + 16);
    }
    const std::size_t na = net.cfg.n_astro, nn = net.cfg.n_neurons();
    for (int i = 0; i < 7; ++i) o.astro[i] = L.add(na * sizeof(double));
    o.ever = L.add(na);
    for (int i = 0; i < 12; ++i) o.neuron[i] = L.add(nn * sizeof(double));
    o.refr = L.add(nn * sizeof(std::int32_t));
    o.spk_n = L.add(nn * sizeof(std::uint32_t));
    const std::size_t R = std::size_t(rp.ring_slots);
    o.wn = ceil_div(nn, 64);
    o.wa = ceil_div(na, 64);
    // Staged bits: the flag words themselves when both fit the budget,
    // otherwise summaries, coarsened until they fit.
    o.split = nn >= kSplitMinNeurons;
    // resident waves per SIMD, set by registers: fused 4 (the cell
    // arithmetic), the split gather kernel 6
    const std::size_t budget = o.split ? lds_budget(kGatherThreads, 24) : lds_budget(rp.block, 16);
    if (std::size_t(o.wn + o.wa) * 8 <= budget) {
        o.gsh_n = o.gsh_s = 0;
        o.sn = 2 * o.wn;
        o.sa = 2 * o.wa;
    } else {
        int g = 6;
        for (;; ++g) {
            o.sn = ceil_div(ceil_div(nn, std::uint64_t(1) << g), 32);
            o.sa = ceil_div(ceil_div(na, std::uint64_t(1) << g), 32);
            if (std::size_t(o.sn + o.sa) * 4 <= budget || g >= 20) break;
        }
        o.gsh_n = o.gsh_s = g;
    }
    o.flag_n = L.add(R * o.wn * 8);
    o.flag_s = L.add(R * o.wa * 8);
    o.val_n = L.add(R * nn * sizeof(double));
    o.val_s = L.add(R * na * sizeof(double));
    o.sum_n = L.add((R + 1) * o.sn * 4);
    o.sum_s = L.add((R + 1) * o.sa * 4);
    o.counters = L.add(kSpikeCounters * sizeof(unsigned long long));
    if (o.split) {
        for (int i = 0; i < 4; ++i) o.in[i] = L.add((i == 3 ? na : nn) * sizeof(double));
        for (int i = 0; i < 4; ++i) o.mask[i] = L.add(std::size_t(i == 3 ? o.wa : o.wn) * 8);
    }
    o.total = L.total;
    return o;
}

using StepFn = void (*)(PlainDev, RunParams, std::uint64_t, PlainStep);
using GatherFn = void (*)(PlainDev, PlainStep);

class PlainEngine final : public Engine {
public:
    PlainEngine(const Network& net, const RunParams& rp, hipStream_t s) : P_(rp), seed_(net.seed) {
        if (rp.d_e != rp.d_i || rp.d_e != rp.d_a2n)
            throw Error(AS_ERR_INVALID, "the plain method needs equal delays for all connection types");
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
        D_.spk_n = (std::uint32_t*)dptr(lay_.spk_n);
        D_.flag_n = (std::uint64_t*)dptr(lay_.flag_n);
        D_.flag_s = (std::uint64_t*)dptr(lay_.flag_s);
        D_.val_n = (double*)dptr(lay_.val_n);
        D_.val_s = (double*)dptr(lay_.val_s);
        D_.sum_n = (std::uint32_t*)dptr(lay_.sum_n);
        D_.sum_s = (std::uint32_t*)dptr(lay_.sum_s);
        D_.counters = (unsigned long long*)dptr(lay_.counters);
        D_.n_astro = n_astro_;
        D_.n_neurons = n_neurons_;
        D_.n_exc = n_exc_;
        D_.wn = lay_.wn;
        D_.wa = lay_.wa;
        D_.sn = lay_.sn;
        D_.sa = lay_.sa;
        D_.gsh_n = lay_.gsh_n;
        D_.gsh_s = lay_.gsh_s;
        D_.astro_waves = ceil_div(n_astro_, 64);
        D_.neuron_waves = ceil_div(n_neurons_, 64);
        lds_bytes_ = std::size_t(lay_.sn + lay_.sa) * 4;

        if (lay_.split) {
            double* in[4];
            std::uint64_t* mk[4];
            for (int i = 0; i < 4; ++i) {
                in[i] = (double*)dptr(lay_.in[i]);
                mk[i] = (std::uint64_t*)dptr(lay_.mask[i]);
            }
            D_.in_exc = in[0]; D_.in_inh = in[1]; D_.in_sic = in[2]; D_.in_e2a = in[3];
            D_.mask_exc = mk[0]; D_.mask_inh = mk[1]; D_.mask_sic = mk[2]; D_.mask_e2a = mk[3];
        }

        // The looping kernel (fused, or the split gather) gets one round of
        // resident workgroups; the split cell kernel one tile per workgroup.
        int dev_id = 0, cus = 0, per_cu = 0;
        AS_HIP_CHECK(hipGetDevice(&dev_id));
        AS_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev_id));
        const bool big = P_.block > 256;
        const void* looping;
        if (lay_.split) {
            gather_ = plain_gather_kernel<kGatherThreads, kLoadsInFlight>;
            cell_ = big ? plain_cell_kernel<1024> : plain_cell_kernel<256>;
            looping = (const void*)gather_;
            looping_threads_ = kGatherThreads;
        } else {
            step_fn_ = big ? plain_step_kernel<1024, kLoadsInFlight, 1> : plain_step_kernel<256, kLoadsInFlight, 4>;
            looping = (const void*)step_fn_;
            looping_threads_ = P_.block;
        }
        AS_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, looping, looping_threads_, lds_bytes_));
        grid_ = std::min<std::uint32_t>(make_tiles(D_.astro_waves, D_.neuron_waves, looping_threads_ / 64).total,
                                        std::uint32_t(std::max(1, per_cu) * std::max(1, cus)));
        cell_grid_ = make_tiles(D_.astro_waves, D_.neuron_waves, P_.block / 64).total;

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
        AS_HIP_CHECK(hipMemsetAsync(D_.spk_n, 0, n_neurons_ * sizeof(std::uint32_t), stream_));
        AS_HIP_CHECK(hipMemsetAsync(D_.ever, 0, n_astro_, stream_));
        // Flags, values and summaries: everything from the ring on is zero at step 0.
        AS_HIP_CHECK(hipMemsetAsync(dptr(lay_.flag_n), 0, lay_.total - lay_.flag_n, stream_));
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
        const int R = P_.ring_slots, R1 = R + 1, d = P_.d_e;
        for (std::int64_t i = 0; i < k; ++i) {
            const std::int64_t step = step_ + i;
            PlainStep S;
            S.step = step;
            S.noise_idx = static_cast<std::uint64_t>(step / P_.steps_per_noise);
            S.sic_refresh = (step - P_.d_a2n >= 0) && ((step - P_.d_a2n) % P_.sic_interval == 0);
            S.rd_n = static_cast<std::uint32_t>((step + R - d) % R);
            S.rd_s = S.rd_n;
            S.wr = static_cast<std::uint32_t>(step % R);
            S.srd_n = static_cast<std::uint32_t>((step + R1 - d) % R1);
            S.srd_s = S.srd_n;
            S.swr = static_cast<std::uint32_t>(step % R1);
            S.sclr = static_cast<std::uint32_t>((step + 1) % R1);
            timing_begin();
            if (lay_.split) {
                hipLaunchKernelGGL(gather_, dim3(grid_), dim3(kGatherThreads), lds_bytes_, stream_, D_, S);
                hipLaunchKernelGGL(cell_, dim3(cell_grid_), dim3(P_.block), 0, stream_, D_, P_, seed_, S);
            } else {
                hipLaunchKernelGGL(step_fn_, dim3(grid_), dim3(P_.block), lds_bytes_, stream_, D_, P_, seed_, S);
            }
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

    void read_spike_counts(std::uint32_t* out) override {
        AS_HIP_CHECK(hipMemcpyAsync(out, D_.spk_n, n_neurons_ * sizeof(std::uint32_t), hipMemcpyDeviceToHost, stream_));
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void read_ever_active(std::uint8_t* out) override {
        AS_HIP_CHECK(hipMemcpyAsync(out, D_.ever, n_astro_, hipMemcpyDeviceToHost, stream_));
        AS_HIP_CHECK(hipStreamSynchronize(stream_));
    }

    void read_raster(std::int64_t first, int n, std::uint8_t* out) override {
        // Slot s % R holds the spike words of step s for the last R steps.
        const int R = P_.ring_slots;
        if (n < 0 || first < 0 || first + n > step_ || first < step_ - R)
            throw Error(AS_ERR_RANGE, "raster steps not held (the plain method keeps the last R steps)");
        std::vector<std::uint64_t> words(lay_.wn);
        for (int i = 0; i < n; ++i) {
            const std::int64_t s = first + i;
            AS_HIP_CHECK(hipMemcpyAsync(words.data(), D_.flag_n + std::size_t(s % R) * lay_.wn, lay_.wn * 8ull,
                                        hipMemcpyDeviceToHost, stream_));
            AS_HIP_CHECK(hipStreamSynchronize(stream_));
            for (index_t j = 0; j < n_neurons_; ++j)
                out[std::size_t(i) * n_neurons_ + j] = static_cast<std::uint8_t>((words[j >> 6] >> (j & 63u)) & 1u);
        }
    }

    SpikeWords spike_words(std::int64_t first, int n) override {
        const int R = P_.ring_slots;
        if (n < 0 || n > kRecordMaxSteps || first < 0 || first + n > step_ || first < step_ - R)
            throw Error(AS_ERR_RANGE, "spike steps not held (the plain method keeps the last R steps)");
        SpikeWords w;
        w.kind = 0;
        w.base = D_.flag_n;
        w.n_neurons = n_neurons_;
        w.n_steps = n;
        for (int i = 0; i < n; ++i) w.off[i] = std::uint64_t((first + i) % R) * lay_.wn;
        return w;
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
    std::size_t lds_bytes_ = 0;
    std::uint32_t grid_ = 1;
    StepFn step_fn_ = nullptr;  // fused
    GatherFn gather_ = nullptr;  // split
    int looping_threads_ = 256;
    std::uint32_t cell_grid_ = 1;
    StepFn cell_ = nullptr;     // split
};

}  // namespace

std::uint64_t plain_footprint(const Network& net, const RunParams& rp) { return plain_layout(net, rp).total; }

std::unique_ptr<Engine> make_plain_engine(const Network& net, const RunParams& rp, hipStream_t s) {
    return std::unique_ptr<Engine>(new PlainEngine(net, rp, s));
}

}  // namespace astrosim
