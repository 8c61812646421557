/* oracle.h -- FP64 host oracle of the astrosim_fused per-step pipeline.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference
 * program (src/astrosim_fused.cu) used to check the HIP implementation and,
 * timed on the host, as the CPU baseline. Nothing in the shipped library
 * links, loads or calls it.
 *
 * It restates, cell by cell and step by step in the reference's two-phase
 * order (all astrocytes, then all neurons, each step):
 *   setup  : counter RNG (src:127-175), sequential HostRng (src:180-216),
 *            scaling law (src:606-614), connectivity generation and CSR
 *            inversion (src:617-735), per-cell parameter draws (src:917-933),
 *            initial state (src:981-1018);
 *   step   : astro_step_kernel (src:374-427) and neuron_step_kernel
 *            (src:433-587) with the host loop scalars (src:1042-1052).
 * Floating point is IEEE double with contraction off; exp/log/cos inside the
 * step come from the shared deterministic as_dmath.h (the setup keeps the C
 * library's log/log1p/sqrt/sin/cos, as the reference's host code does).
 *
 * Pinning (each check lives in tests/): splitmix64 known answers, exact
 * uniform identity, Knuth-Poisson brute force and moments, Box-Muller pair
 * identities, scaling-law closed forms, connectivity invariants (no
 * autapses, row order, pool and out-degree caps, in-degree statistics),
 * serial == threaded bitwise; and, for the step arithmetic
 * (tests/test_oracle_physics.py), through the per-cell entry points below:
 * Li-Rinzel values at degenerate points and its fixed point, RK4's fourth
 * order, the clamps, the AdEx values at threshold, at rest and above the
 * peak, the alpha-cascade peak, the refractory clamp, the STP first spike
 * and recursion, and zero-term invariance of the gathers.
 */
#ifndef ASTROSIM_ORACLE_H
#define ASTROSIM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_net orc_net;
typedef struct orc_sim orc_sim;

/* Counter RNG primitives (src:127-175), exported for known-answer tests. */
uint64_t orc_rng_bits(uint64_t seed, uint64_t stream, uint64_t counter);
double orc_rng_uniform(uint64_t seed, uint64_t stream, uint64_t counter);
double orc_rng_normal(uint64_t seed, uint64_t stream);
int orc_rng_poisson(uint64_t seed, uint64_t stream, double lambda);
/* HostRng sequence (src:180-208): n successive normal() values. */
void orc_hostrng_normals(uint64_t seed, uint64_t stream, int64_t n, double* out);
/* dmath entry points, exported for accuracy tests. */
double orc_exp(double x);
double orc_log(double x);
double orc_cos(double x);

/* ---- network (src:595-735, 917-933) ---- */
orc_net* orc_net_build(uint32_t n_total, uint64_t seed);
/* Wraps connectivity and parameters produced elsewhere (copied), so the
 * stepper can be timed on a network too large for orc_net_build's
 * pair-buffer generation. csr_rs/csr_src: exc, inh, sic, e2a; params: Ca_tot,
 * IP3_0, tau_IP3, delta_IP3, V_reset, b. */
orc_net* orc_net_from_arrays(uint32_t n_total, uint64_t seed, const uint32_t* const csr_rs[4],
                             const uint32_t* const csr_src[4], const double* const params[6]);
void orc_net_free(orc_net* net);
/* counts[0..3] = n_exc, n_inh, n_astro, n_neurons; p = p_primary */
void orc_net_sizes(const orc_net* net, uint32_t counts[4], double* p_primary);
/* which: 0 exc, 1 inh, 2 sic, 3 e2a. Returns the edge count; copies when the
 * pointers are non-null (row_start has n_targets+1 entries). */
uint64_t orc_net_csr(const orc_net* net, int which, uint32_t* row_start, uint32_t* src);
/* which: 0 Ca_tot, 1 IP3_0, 2 tau_IP3, 3 delta_IP3 (n_astro), 4 V_reset, 5 b (n_neurons) */
void orc_net_param(const orc_net* net, int which, double* out);
double orc_net_build_seconds(const orc_net* net);

/* ---- simulation ---- */
typedef struct orc_opts {
    double dt;         /* --dt */
    int substeps;      /* --substeps */
    double w_n2a;      /* NaN keeps the default 0.2 */
    double w_a2n;      /* NaN keeps the default 1.0 */
    int stp_enabled;   /* 0 for --no-stp */
    int threads;       /* OpenMP threads for the step (<= 0: all) */
} orc_opts;

orc_sim* orc_sim_create(const orc_net* net, const orc_opts* opts);
void orc_sim_free(orc_sim* sim);
/* Advances k steps. If raster != NULL it receives k*n_neurons bytes, 1 where
 * the neuron spiked in that step. */
void orc_sim_step(orc_sim* sim, int64_t k, uint8_t* raster);
int64_t orc_sim_steps_done(const orc_sim* sim);
uint64_t orc_sim_spike_count(const orc_sim* sim);
void orc_sim_reset_ever_active(orc_sim* sim);
/* field ids: 0 Ca 1 IP3 2 h (n_astro); 3 V 4 w 5 g_ex 6 dg_ex 7 g_in 8 dg_in
 * 9 I_sic_hold 10 stp_x 11 stp_u 12 stp_t_last (n_neurons) */
void orc_sim_read(const orc_sim* sim, int field, double* out);
void orc_sim_read_refractory(const orc_sim* sim, int32_t* out);
void orc_sim_read_ever_active(const orc_sim* sim, uint8_t* out);

/* ---- per-cell step arithmetic, exported for known-answer tests ---- */
/* Li-Rinzel right-hand side (src:236-253): out = {dCa, dIP3, dh}. */
void orc_astro_rhs(double Ca_tot, double IP3_0, double tau_IP3, double Ca, double IP3, double h, double noise,
                   double out[3]);
/* IP3 jump, `substeps` RK4 steps of h_step and clamps (src:255-278);
 * state = {Ca, IP3, h}, updated in place. */
void orc_astro_advance(double Ca_tot, double IP3_0, double tau_IP3, double delta_IP3, double ip3_input,
                       double noise, double h_step, int substeps, double state[3]);
/* SIC factor of a calcium level (src:416-417). */
double orc_sic_factor(double Ca);
/* AdEx right-hand side of the excitatory or inhibitory population
 * (src:281-290): out = {dV, dw}. */
void orc_neuron_rhs(int excitatory, double V, double w, double g_ex, double g_in, double I_ext, double out[2]);
/* Substep loop (src:508-553); state = {V, w, g_ex, dg_ex, g_in, dg_in}
 * updated in place. Returns 1 on a spike. */
int orc_neuron_substeps(int excitatory, double state[6], int32_t* refractory, double I_ext, double V_reset,
                        double b, double h_step, int substeps);
/* e / tau_syn of the population's excitatory (0) or inhibitory (1) synapses (src:944-951). */
double orc_psc_init(int excitatory, int inhibitory_synapse);
/* STP on a spike at t_now (src:565-576); state = {x, u, t_last} in place;
 * returns x*u. */
double orc_stp_on_spike(double state[3], double t_now);
/* Sum of ring_row over the CSR row of `target`, in row order. */
double orc_gather(const uint32_t* row_start, const uint32_t* src, uint32_t n_targets, const double* ring_row,
                  uint32_t target);
/* {d_e, d_i, d_a2n, ring depth R, steps_per_noise} of a simulation (src:954-957, 1043). */
void orc_sim_params(const orc_sim* sim, int32_t out[5]);

#ifdef __cplusplus
}
#endif

#endif
