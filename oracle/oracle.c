/* oracle.c -- FP64 host oracle (test infrastructure only; see oracle.h).
 *
 * Every function cites the reference span it restates. Build with
 * -ffp-contract=off (oracle/Makefile does).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "as_dmath.h"

/* ======================================================================
 * Counter RNG (src/astrosim_fused.cu:127-175)
 * ====================================================================== */
uint64_t orc_rng_bits(uint64_t seed, uint64_t stream, uint64_t counter) {
    uint64_t z = seed;
    z += stream * 0x9E3779B97F4A7C15ULL;
    z += counter * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

double orc_rng_uniform(uint64_t seed, uint64_t stream, uint64_t counter) {
    return (double)(orc_rng_bits(seed, stream, counter) >> 11) * (1.0 / 9007199254740992.0);
}

double orc_rng_normal(uint64_t seed, uint64_t stream) {
    double u1 = orc_rng_uniform(seed, stream, 0);
    if (u1 <= 1e-300) u1 = 1e-300;
    double u2 = orc_rng_uniform(seed, stream, 1);
    return sqrt(-2.0 * as_log(u1)) * as_cos(6.283185307179586476925286766559 * u2);
}

int orc_rng_poisson(uint64_t seed, uint64_t stream, double lambda) {
    if (lambda <= 0.0) return 0;
    uint64_t c = 0;
    if (lambda > 30.0) {
        double u1 = orc_rng_uniform(seed, stream, c++);
        if (u1 <= 1e-300) u1 = 1e-300;
        double u2 = orc_rng_uniform(seed, stream, c++);
        double z = sqrt(-2.0 * as_log(u1)) * as_cos(6.283185307179586476925286766559 * u2);
        double v = lambda + sqrt(lambda) * z;
        return v < 0.0 ? 0 : (int)(v + 0.5);
    }
    double limit = as_exp(-lambda);
    double product = orc_rng_uniform(seed, stream, c++);
    int count = 0;
    while (product > limit) {
        ++count;
        product *= orc_rng_uniform(seed, stream, c++);
    }
    return count;
}

double orc_exp(double x) { return as_exp(x); }
double orc_log(double x) { return as_log(x); }
double orc_cos(double x) { return as_cos(x); }

/* ======================================================================
 * Sequential host RNG (src/astrosim_fused.cu:180-216)
 * ====================================================================== */
typedef struct {
    uint64_t seed, stream, counter;
    int has_spare;
    double spare;
} hrng;

static void hrng_init(hrng* r, uint64_t seed, uint64_t stream) {
    r->seed = seed; r->stream = stream; r->counter = 0; r->has_spare = 0; r->spare = 0.0;
}
static double hrng_uniform(hrng* r) { return orc_rng_uniform(r->seed, r->stream, r->counter++); }
static double hrng_normal(hrng* r) {
    if (r->has_spare) { r->has_spare = 0; return r->spare; }
    double u1;
    do { u1 = hrng_uniform(r); } while (u1 <= 1e-300);
    double u2 = hrng_uniform(r);
    double rad = sqrt(-2.0 * log(u1));
    double theta = 6.283185307179586476925286766559 * u2;
    /* one sincos call, as an optimising host compiler emits for the
     * reference's sin(theta)/cos(theta) pair */
    double sn, cs;
    sincos(theta, &sn, &cs);
    r->spare = rad * sn;
    r->has_spare = 1;
    return rad * cs;
}
static double hrng_normal_redraw(hrng* r, double mean, double sd, double lo, double hi) {
    for (int i = 0; i < 1000; ++i) {
        double v = mean + sd * hrng_normal(r);
        if (v >= lo && v <= hi) return v;
    }
    double m = mean > lo ? mean : lo; /* std::max(mean, lo) */
    return hi < m ? hi : m;            /* std::min(.., hi) */
}
static double draw_scaled(double mean, int enabled, double lower, double upper, double var, hrng* r) {
    if (!enabled || var <= 0.0) return mean;
    double lo = mean < 0.0 ? mean * upper : mean * lower;
    double hi = mean < 0.0 ? mean * lower : mean * upper;
    return hrng_normal_redraw(r, mean, var * fabs(mean), lo, hi);
}

void orc_hostrng_normals(uint64_t seed, uint64_t stream, int64_t n, double* out) {
    hrng r;
    hrng_init(&r, seed, stream);
    for (int64_t i = 0; i < n; ++i) out[i] = hrng_normal(&r);
}

/* ======================================================================
 * Biology constants (src/astrosim_fused.cu:742-781)
 * ====================================================================== */
#define B_CA_TOT 1.8958264782990153
#define B_IP3_0 0.010925583871830699
#define B_TAU_IP3 1058.7589568400845
#define B_DELTA_IP3 0.04911592657233997
#define B_SIC_SCALE 1.0
#define B_KD_IP3_1 0.13
#define B_KD_IP3_2 0.9434
#define B_KD_ACT 0.08234
#define B_KD_INH 1.049
#define B_KM_SERCA 0.1
#define B_SIC_TH 0.19669
#define B_K_IP3R 0.0002
#define B_RATE_IP3R 0.006
#define B_RATE_L 0.00011
#define B_RATE_SERCA 0.0009
#define B_RATIO_ER_CYT 0.185
#define B_CA_INIT 0.073
#define B_H_INIT 0.793
#define B_IP3_INIT 0.010925583871830699

typedef struct {
    double C_m, g_L, E_L, V_th, Delta_T, a, tau_w, V_peak, t_ref, E_ex, E_in, tau_syn_ex,
        tau_syn_in, I_e, psc_ex, psc_in, V_m, V_reset, b;
} nrn_par;

static nrn_par exc_par(void) {
    nrn_par p = {130.0, 18.0, -58.0, -50.0, 2.0, 4.0, 450.0, 0.0, 0.0, 0.0, -85.0, 0.2, 2.0,
                 0.0, 0.0, 0.0, -50.0, -50.0, 300.0};
    p.psc_ex = 2.7182818284590452353602874713527 / p.tau_syn_ex;
    p.psc_in = 2.7182818284590452353602874713527 / p.tau_syn_in;
    return p;
}
static nrn_par inh_par(void) {
    nrn_par p = {104.0, 4.3, -65.0, -52.0, 2.0, -0.8, 264.0, 0.0, 0.0, 0.0, -85.0, 0.2, 2.0,
                 0.0, 0.0, 0.0, -60.0, -60.0, 130.0};
    p.psc_ex = 2.7182818284590452353602874713527 / p.tau_syn_ex;
    p.psc_in = 2.7182818284590452353602874713527 / p.tau_syn_in;
    return p;
}

/* ======================================================================
 * Network (src/astrosim_fused.cu:595-735, 917-933)
 * ====================================================================== */
typedef struct { uint32_t* row_start; uint32_t* src; uint32_t n_targets; uint64_t n_edges; } csr_t;

struct orc_net {
    uint32_t n_total, n_exc, n_inh, n_astro, n_neurons;
    double p_primary;
    uint64_t seed;
    csr_t csr[4]; /* exc, inh, sic, e2a */
    double *Ca_tot, *IP3_0, *tau_IP3, *delta_IP3, *V_reset, *b;
    double build_seconds;
};

typedef struct { uint32_t* s; uint32_t* d; uint64_t n, cap; } pairs_t;
static void pairs_push(pairs_t* p, uint32_t s, uint32_t d) {
    if (p->n == p->cap) {
        p->cap = p->cap ? p->cap * 2 : 1024;
        p->s = (uint32_t*)realloc(p->s, p->cap * sizeof(uint32_t));
        p->d = (uint32_t*)realloc(p->d, p->cap * sizeof(uint32_t));
    }
    p->s[p->n] = s; p->d[p->n] = d; p->n++;
}

/* Stable counting sort by target (src/astrosim_fused.cu:307-314). */
static void csr_build(csr_t* c, uint32_t n_targets, const pairs_t* e) {
    c->n_targets = n_targets;
    c->n_edges = e->n;
    c->row_start = (uint32_t*)calloc((size_t)n_targets + 1, sizeof(uint32_t));
    for (uint64_t k = 0; k < e->n; ++k) c->row_start[e->d[k] + 1]++;
    for (uint32_t i = 0; i < n_targets; ++i) c->row_start[i + 1] += c->row_start[i];
    c->src = (uint32_t*)malloc((e->n ? e->n : 1) * sizeof(uint32_t));
    uint32_t* cur = (uint32_t*)malloc(((size_t)n_targets + 1) * sizeof(uint32_t));
    memcpy(cur, c->row_start, (size_t)n_targets * sizeof(uint32_t));
    for (uint64_t k = 0; k < e->n; ++k) c->src[cur[e->d[k]]++] = e->s[k];
    free(cur);
}

/* src/astrosim_fused.cu:617-624 */
static uint32_t geometric_skip(double p, uint32_t remaining, hrng* r) {
    if (remaining == 0 || p >= 1.0) return 0;
    if (p <= 0.0) return remaining;
    double u = hrng_uniform(r);
    if (u >= 1.0) u = nextafter(1.0, 0.0);
    double g = floor(log1p(-u) / log1p(-p));
    return (g >= (double)remaining) ? remaining : (uint32_t)g;
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

orc_net* orc_net_build(uint32_t n_total, uint64_t seed) {
    if (n_total == 0) return NULL;
    orc_net* net = (orc_net*)calloc(1, sizeof(orc_net));
    /* ScalingConfig::make (src:606-614) */
    net->n_total = n_total;
    long ne = lround(n_total * 0.8);
    net->n_exc = (uint32_t)ne < 1 ? 1 : (uint32_t)ne;
    uint32_t ni = n_total - net->n_exc;
    net->n_inh = ni < 1 ? 1 : ni;
    long na = lround(n_total / 5.0);
    net->n_astro = (uint32_t)na < 1 ? 1 : (uint32_t)na;
    net->p_primary = 80.0 / (double)net->n_exc;
    if (net->p_primary > 1.0) net->p_primary = 1.0;
    net->n_neurons = net->n_exc + net->n_inh;
    net->seed = seed;
    const uint32_t n_exc = net->n_exc, n_inh = net->n_inh, n_astro = net->n_astro;
    const uint32_t n_neurons = net->n_neurons;

    double t0 = now_s();
    hrng r;
    hrng_init(&r, seed, 0x5EEDULL);
    for (uint32_t i = 0; i < n_astro; ++i) { hrng_normal(&r); hrng_normal(&r); hrng_normal(&r); hrng_normal(&r); }
    for (uint32_t i = 0; i < n_neurons; ++i) { hrng_normal(&r); hrng_normal(&r); }

    /* pools (src:651-671) */
    int take = n_astro < 5 ? (int)n_astro : 5;
    uint32_t* pool = (uint32_t*)malloc((size_t)n_neurons * take * sizeof(uint32_t));
    uint32_t* cand = (uint32_t*)malloc((size_t)n_astro * sizeof(uint32_t));
    for (uint32_t post = 0; post < n_neurons; ++post) {
        uint32_t* pl = pool + (size_t)post * take;
        if ((int64_t)take * 4 >= (int64_t)n_astro) {
            size_t size = n_astro;
            for (uint32_t i = 0; i < n_astro; ++i) cand[i] = i;
            for (int k = 0; k < take; ++k) {
                size_t pick = (size_t)(hrng_uniform(&r) * (double)(size - (size_t)k)) + (size_t)k;
                size_t j = pick < size - 1 ? pick : size - 1;
                uint32_t tmp = cand[k]; cand[k] = cand[j]; cand[j] = tmp;
                pl[k] = cand[k];
            }
        } else {
            int have = 0;
            while (have < take) {
                uint32_t c = (uint32_t)(hrng_uniform(&r) * (double)n_astro) % n_astro;
                int dup = 0;
                for (int q = 0; q < have; ++q) if (pl[q] == c) { dup = 1; break; }
                if (!dup) pl[have++] = c;
            }
        }
    }
    free(cand);

    uint32_t* out_deg = (uint32_t*)calloc(n_astro, sizeof(uint32_t));
    pairs_t e_exc = {0}, e_inh = {0}, e_e2a = {0}, e_sic = {0};
    const double p = net->p_primary;
    /* connect_block (src:688-710), called for posts [0,n_exc) then [n_exc,n) */
    for (int blk = 0; blk < 2; ++blk) {
        uint32_t off = blk == 0 ? 0 : n_exc;
        uint32_t cnt = blk == 0 ? n_exc : n_inh;
        for (uint32_t pre = 0; pre < n_exc; ++pre) {
            uint32_t pos = 0;
            while (pos < cnt) {
                pos += geometric_skip(p, cnt - pos, &r);
                if (pos >= cnt) break;
                uint32_t post = off + pos;
                ++pos;
                if (pre == post) continue;
                pairs_push(&e_exc, pre, post);
                if (n_astro == 0 || hrng_uniform(&r) >= 0.2) continue;
                const uint32_t* pl = pool + (size_t)post * take;
                uint32_t astro = pl[(size_t)(hrng_uniform(&r) * (double)take) % (size_t)take];
                if (out_deg[astro] >= 10000u) continue;
                pairs_push(&e_e2a, pre, astro);
                pairs_push(&e_sic, astro, post);
                ++out_deg[astro];
            }
        }
    }
    /* inhibitory rows, local source ids (src:712-723) */
    for (uint32_t pre = 0; pre < n_inh; ++pre) {
        uint32_t pre_global = n_exc + pre;
        uint32_t pos = 0;
        while (pos < n_neurons) {
            pos += geometric_skip(p, n_neurons - pos, &r);
            if (pos >= n_neurons) break;
            uint32_t post = pos;
            ++pos;
            if (pre_global == post) continue;
            pairs_push(&e_inh, pre, post);
        }
    }
    csr_build(&net->csr[0], n_neurons, &e_exc);
    csr_build(&net->csr[1], n_neurons, &e_inh);
    csr_build(&net->csr[2], n_neurons, &e_sic);
    csr_build(&net->csr[3], n_astro, &e_e2a);
    free(e_exc.s); free(e_exc.d); free(e_inh.s); free(e_inh.d);
    free(e_e2a.s); free(e_e2a.d); free(e_sic.s); free(e_sic.d);
    free(out_deg); free(pool);
    net->build_seconds = now_s() - t0;

    /* per-cell parameters (src:917-933) */
    hrng pr;
    hrng_init(&pr, seed, 0x5EEDULL);
    net->Ca_tot = (double*)malloc(n_astro * sizeof(double));
    net->IP3_0 = (double*)malloc(n_astro * sizeof(double));
    net->tau_IP3 = (double*)malloc(n_astro * sizeof(double));
    net->delta_IP3 = (double*)malloc(n_astro * sizeof(double));
    for (uint32_t i = 0; i < n_astro; ++i) {
        net->Ca_tot[i] = draw_scaled(B_CA_TOT, 1, 0.9, 1.05, 0.1, &pr);
        net->IP3_0[i] = draw_scaled(B_IP3_0, 1, 0.9, 1.05, 0.1, &pr);
        net->tau_IP3[i] = draw_scaled(B_TAU_IP3, 1, 0.9, 1.05, 0.1, &pr);
        net->delta_IP3[i] = draw_scaled(B_DELTA_IP3, 1, 0.9, 1.05, 0.1, &pr);
    }
    net->V_reset = (double*)malloc(n_neurons * sizeof(double));
    net->b = (double*)malloc(n_neurons * sizeof(double));
    nrn_par pe = exc_par(), pi = inh_par();
    for (uint32_t i = 0; i < n_neurons; ++i) {
        int exc = i < n_exc;
        net->V_reset[i] = draw_scaled(exc ? pe.V_reset : pi.V_reset, 1, 0.9, 1.1, 0.05, &pr);
        net->b[i] = draw_scaled(exc ? pe.b : pi.b, 1, 0.9, 1.1, 0.05, &pr);
    }
    return net;
}

orc_net* orc_net_from_arrays(uint32_t n_total, uint64_t seed, const uint32_t* const csr_rs[4],
                             const uint32_t* const csr_src[4], const double* const params[6]) {
    if (n_total == 0) return NULL;
    orc_net* net = (orc_net*)calloc(1, sizeof(orc_net));
    net->n_total = n_total;
    long ne = lround(n_total * 0.8);
    net->n_exc = (uint32_t)ne < 1 ? 1 : (uint32_t)ne;
    uint32_t ni = n_total - net->n_exc;
    net->n_inh = ni < 1 ? 1 : ni;
    long na = lround(n_total / 5.0);
    net->n_astro = (uint32_t)na < 1 ? 1 : (uint32_t)na;
    net->p_primary = 80.0 / (double)net->n_exc;
    if (net->p_primary > 1.0) net->p_primary = 1.0;
    net->n_neurons = net->n_exc + net->n_inh;
    net->seed = seed;
    for (int i = 0; i < 4; ++i) {
        uint32_t nt = i == 3 ? net->n_astro : net->n_neurons;
        csr_t* c = &net->csr[i];
        c->n_targets = nt;
        c->n_edges = csr_rs[i][nt];
        c->row_start = (uint32_t*)malloc(((size_t)nt + 1) * sizeof(uint32_t));
        memcpy(c->row_start, csr_rs[i], ((size_t)nt + 1) * sizeof(uint32_t));
        c->src = (uint32_t*)malloc((c->n_edges ? c->n_edges : 1) * sizeof(uint32_t));
        if (c->n_edges) memcpy(c->src, csr_src[i], c->n_edges * sizeof(uint32_t));
    }
    double** dst[6] = {&net->Ca_tot, &net->IP3_0, &net->tau_IP3, &net->delta_IP3, &net->V_reset, &net->b};
    for (int i = 0; i < 6; ++i) {
        size_t n = i < 4 ? net->n_astro : net->n_neurons;
        *dst[i] = (double*)malloc(n * sizeof(double));
        memcpy(*dst[i], params[i], n * sizeof(double));
    }
    return net;
}

void orc_net_free(orc_net* net) {
    if (!net) return;
    for (int i = 0; i < 4; ++i) { free(net->csr[i].row_start); free(net->csr[i].src); }
    free(net->Ca_tot); free(net->IP3_0); free(net->tau_IP3); free(net->delta_IP3);
    free(net->V_reset); free(net->b);
    free(net);
}

void orc_net_sizes(const orc_net* net, uint32_t counts[4], double* p_primary) {
    counts[0] = net->n_exc; counts[1] = net->n_inh; counts[2] = net->n_astro; counts[3] = net->n_neurons;
    if (p_primary) *p_primary = net->p_primary;
}

uint64_t orc_net_csr(const orc_net* net, int which, uint32_t* row_start, uint32_t* src) {
    const csr_t* c = &net->csr[which];
    if (row_start) memcpy(row_start, c->row_start, ((size_t)c->n_targets + 1) * sizeof(uint32_t));
    if (src && c->n_edges) memcpy(src, c->src, c->n_edges * sizeof(uint32_t));
    return c->n_edges;
}

void orc_net_param(const orc_net* net, int which, double* out) {
    const double* srcs[6] = {net->Ca_tot, net->IP3_0, net->tau_IP3, net->delta_IP3, net->V_reset, net->b};
    size_t n = which < 4 ? net->n_astro : net->n_neurons;
    memcpy(out, srcs[which], n * sizeof(double));
}

double orc_net_build_seconds(const orc_net* net) { return net->build_seconds; }

/* ======================================================================
 * Simulation state and the two per-step phases
 * ====================================================================== */
struct orc_sim {
    const orc_net* net;
    double dt, h_step;
    int substeps, stp_enabled, threads;
    double w_e, w_i, w_n2a, w_a2n;
    int d_e, d_i, d_a2n, R, steps_per_noise;
    int64_t step;
    nrn_par pe, pi;
    double *Ca, *IP3, *h;
    double *V, *w, *g_ex, *dg_ex, *g_in, *dg_in, *I_sic_hold;
    int32_t* refr;
    double *stp_x, *stp_u, *stp_t_last;
    double *ring_neuron, *ring_sic; /* [R][cells] */
    uint8_t* ever_active;
    uint64_t spikes;
};

static int delay_to_steps(double delay_ms, double dt) {
    int s = (int)(delay_ms / dt + 0.5);
    return s < 1 ? 1 : s;
}

orc_sim* orc_sim_create(const orc_net* net, const orc_opts* o) {
    orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
    s->net = net;
    s->dt = o->dt;
    s->substeps = o->substeps;
    s->h_step = o->dt / (double)o->substeps;
    s->stp_enabled = o->stp_enabled;
    s->threads = o->threads;
    s->w_e = 5.0; s->w_i = -5.0;
    s->w_n2a = isnan(o->w_n2a) ? 0.2 : o->w_n2a;
    s->w_a2n = isnan(o->w_a2n) ? 1.0 : o->w_a2n;
    /* src:954-957 */
    s->d_e = delay_to_steps(1.0, o->dt);
    s->d_i = delay_to_steps(1.0, o->dt);
    s->d_a2n = delay_to_steps(1.0, o->dt);
    int m = s->d_e > s->d_i ? s->d_e : s->d_i;
    m = m > s->d_a2n ? m : s->d_a2n;
    s->R = m + 1;
    /* noise_dt = 10*dt (src:848), steps_per_noise (src:1043) */
    double noise_dt = 10.0 * o->dt;
    long long spn = llround(noise_dt / o->dt);
    s->steps_per_noise = spn < 1 ? 1 : (int)spn;
    s->pe = exc_par();
    s->pi = inh_par();
    size_t na = net->n_astro, nn = net->n_neurons;
    s->Ca = (double*)malloc(na * sizeof(double));
    s->IP3 = (double*)malloc(na * sizeof(double));
    s->h = (double*)malloc(na * sizeof(double));
    for (size_t i = 0; i < na; ++i) { s->Ca[i] = B_CA_INIT; s->IP3[i] = B_IP3_INIT; s->h[i] = B_H_INIT; }
    s->V = (double*)malloc(nn * sizeof(double));
    for (size_t i = 0; i < nn; ++i) s->V[i] = i < net->n_exc ? s->pe.V_m : s->pi.V_m;
    s->w = (double*)calloc(nn, sizeof(double));
    s->g_ex = (double*)calloc(nn, sizeof(double));
    s->dg_ex = (double*)calloc(nn, sizeof(double));
    s->g_in = (double*)calloc(nn, sizeof(double));
    s->dg_in = (double*)calloc(nn, sizeof(double));
    s->I_sic_hold = (double*)calloc(nn, sizeof(double));
    s->refr = (int32_t*)calloc(nn, sizeof(int32_t));
    s->stp_x = (double*)malloc(nn * sizeof(double));
    s->stp_u = (double*)malloc(nn * sizeof(double));
    s->stp_t_last = (double*)malloc(nn * sizeof(double));
    for (size_t i = 0; i < nn; ++i) { s->stp_x[i] = 1.0; s->stp_u[i] = 0.5; s->stp_t_last[i] = -1e12; }
    s->ring_neuron = (double*)calloc((size_t)s->R * nn, sizeof(double));
    s->ring_sic = (double*)calloc((size_t)s->R * na, sizeof(double));
    s->ever_active = (uint8_t*)calloc(na, 1);
    return s;
}

void orc_sim_free(orc_sim* s) {
    if (!s) return;
    free(s->Ca); free(s->IP3); free(s->h);
    free(s->V); free(s->w); free(s->g_ex); free(s->dg_ex); free(s->g_in); free(s->dg_in);
    free(s->I_sic_hold); free(s->refr); free(s->stp_x); free(s->stp_u); free(s->stp_t_last);
    free(s->ring_neuron); free(s->ring_sic); free(s->ever_active);
    free(s);
}

/* Li-Rinzel RHS (src:236-253) */
static inline void astro_rhs(double Ca_tot, double IP3_0, double tau_IP3, double Ca, double IP3,
                             double h, double noise, double* dCa, double* dIP3, double* dh) {
    double Ca_ER = (Ca_tot - Ca) / B_RATIO_ER_CYT;
    double m_inf = IP3 / (IP3 + B_KD_IP3_1);
    double n_inf = Ca / (Ca + B_KD_ACT);
    double m3 = m_inf * m_inf * m_inf;
    double n3 = n_inf * n_inf * n_inf;
    double h3 = h * h * h;
    double J_channel = B_RATIO_ER_CYT * B_RATE_IP3R * m3 * n3 * h3 * (Ca_ER - Ca);
    double J_pump = B_RATE_SERCA * Ca * Ca / (B_KM_SERCA * B_KM_SERCA + Ca * Ca);
    double J_leak = B_RATIO_ER_CYT * B_RATE_L * (Ca_ER - Ca);
    double alpha_h = B_K_IP3R * B_KD_INH * (IP3 + B_KD_IP3_1) / (IP3 + B_KD_IP3_2);
    double beta_h = B_K_IP3R * Ca;
    *dCa = J_channel - J_pump + J_leak + noise;
    *dIP3 = (IP3_0 - IP3) / tau_IP3;
    *dh = alpha_h * (1.0 - h) - beta_h * h;
}

/* AdEx RHS (src:281-290) */
static inline void neuron_rhs(const nrn_par* p, double V, double w, double g_ex, double g_in,
                              double I_ext, double* dV, double* dw) {
    double Vc = V < p->V_peak ? V : p->V_peak;
    double I_spike = p->g_L * p->Delta_T * as_exp((Vc - p->V_th) / p->Delta_T);
    double I_syn_ex = g_ex * (Vc - p->E_ex);
    double I_syn_in = g_in * (Vc - p->E_in);
    *dV = (-p->g_L * (Vc - p->E_L) + I_spike - I_syn_ex - I_syn_in - w + p->I_e + I_ext) / p->C_m;
    *dw = (p->a * (Vc - p->E_L) - w) / p->tau_w;
}

/* astro_advance (src:255-278): IP3 jump, then `substeps` classic RK4 steps
 * of size hs, each followed by the clamps. */
static void astro_advance(double Ca_tot, double IP3_0, double tau, double delta_IP3, double ip3_input,
                          double noise, double hs, int substeps, double* Ca, double* IP3, double* h) {
    double ca = *Ca, ip3 = *IP3, hh = *h;
    if (ip3_input != 0.0) ip3 += delta_IP3 * ip3_input;
    for (int sub = 0; sub < substeps; ++sub) {
        double k1C, k1I, k1h, k2C, k2I, k2h, k3C, k3I, k3h, k4C, k4I, k4h;
        astro_rhs(Ca_tot, IP3_0, tau, ca, ip3, hh, noise, &k1C, &k1I, &k1h);
        astro_rhs(Ca_tot, IP3_0, tau, ca + 0.5 * hs * k1C, ip3 + 0.5 * hs * k1I, hh + 0.5 * hs * k1h,
                  noise, &k2C, &k2I, &k2h);
        astro_rhs(Ca_tot, IP3_0, tau, ca + 0.5 * hs * k2C, ip3 + 0.5 * hs * k2I, hh + 0.5 * hs * k2h,
                  noise, &k3C, &k3I, &k3h);
        astro_rhs(Ca_tot, IP3_0, tau, ca + hs * k3C, ip3 + hs * k3I, hh + hs * k3h, noise, &k4C,
                  &k4I, &k4h);
        ca += (hs / 6.0) * (k1C + 2.0 * k2C + 2.0 * k3C + k4C);
        ip3 += (hs / 6.0) * (k1I + 2.0 * k2I + 2.0 * k3I + k4I);
        hh += (hs / 6.0) * (k1h + 2.0 * k2h + 2.0 * k3h + k4h);
        ca = ca < 0.0 ? 0.0 : (ca > Ca_tot ? Ca_tot : ca);
        hh = hh < 0.0 ? 0.0 : (hh > 1.0 ? 1.0 : hh);
        ip3 = ip3 < 0.0 ? 0.0 : ip3;
    }
    *Ca = ca; *IP3 = ip3; *h = hh;
}

/* SIC factor (src:416-417) */
static double sic_factor(double ca) {
    double y = (ca - B_SIC_TH) * 1000.0;
    return (y <= 1.0) ? 0.0 : B_SIC_SCALE * as_log(y);
}

/* A gather: sum of the ring row's values over the CSR row of `target`, in
 * row order (src:384-390, 450-463, 482-490). */
static double gather_sum(const csr_t* c, const double* row, uint32_t target) {
    double sum = 0.0;
    for (uint32_t k = c->row_start[target]; k < c->row_start[target + 1]; ++k) sum += row[c->src[k]];
    return sum;
}

/* astro_step_kernel body for astrocyte a (src:374-427) */
static void astro_cell(orc_sim* s, uint32_t a, int64_t step) {
    const orc_net* net = s->net;
    const csr_t* c = &net->csr[3];
    int read_slot = (int)((step + s->R - s->d_e) % s->R);
    const double* row = s->ring_neuron + (size_t)read_slot * net->n_neurons;
    double gathered = gather_sum(c, row, a);
    double ip3_input = s->w_n2a * gathered;
    const double rate = 3.029715502945856;
    if (rate > 0.0) {
        double lambda = rate * s->dt * 1e-3;
        int ev = orc_rng_poisson(net->seed ^ 0xC2B2AE3D27D4EB4FULL, (uint64_t)step * 1000003ULL + a, lambda);
        if (ev > 0) ip3_input += 1.0 * (double)ev;
    }
    double noise = 0.0003 * orc_rng_normal(net->seed ^ 0x9E3779B97F4A7C15ULL, (uint64_t)step * 1000003ULL + a);
    double ca = s->Ca[a], ip3 = s->IP3[a], hh = s->h[a];
    astro_advance(net->Ca_tot[a], net->IP3_0[a], net->tau_IP3[a], net->delta_IP3[a], ip3_input, noise, s->h_step,
                  s->substeps, &ca, &ip3, &hh);
    s->Ca[a] = ca; s->IP3[a] = ip3; s->h[a] = hh;
    double factor = sic_factor(ca);
    if (factor > 0.0) s->ever_active[a] = 1;
    s->ring_sic[(size_t)(step % s->R) * net->n_astro + a] = factor;
}

/* One neuron's substep loop once its inputs are known (src:508-553): the
 * exact alpha-cascade propagator, the refractory clamp (Euler on w only),
 * RK4 on (V, w) with the conductances at t, the midpoint and t+h, spike
 * detection and reset. st = {V, w, g_ex, dg_ex, g_in, dg_in}. Returns 1 on
 * a spike. */
static int neuron_substeps(const nrn_par* p, double* st, int32_t* refr, double I_ext, double V_reset, double b,
                           double hs, int substeps) {
    double V = st[0], w = st[1], g_ex = st[2], dg_ex = st[3], g_in = st[4], dg_in = st[5];
    int refractory = *refr;
    int spiked = 0;
    for (int sub = 0; sub < substeps; ++sub) {
        double ex_decay = as_exp(-hs / p->tau_syn_ex);
        double in_decay = as_exp(-hs / p->tau_syn_in);
        double g_ex_next = ex_decay * (g_ex + hs * dg_ex);
        double dg_ex_next = ex_decay * dg_ex;
        double g_in_next = in_decay * (g_in + hs * dg_in);
        double dg_in_next = in_decay * dg_in;
        if (refractory > 0) {
            V = V_reset;
            double dV, dw;
            neuron_rhs(p, V, w, g_ex, g_in, I_ext, &dV, &dw);
            w += hs * dw;
        } else {
            double g_ex_mid = 0.5 * (g_ex + g_ex_next);
            double g_in_mid = 0.5 * (g_in + g_in_next);
            double k1V, k1w, k2V, k2w, k3V, k3w, k4V, k4w;
            neuron_rhs(p, V, w, g_ex, g_in, I_ext, &k1V, &k1w);
            neuron_rhs(p, V + 0.5 * hs * k1V, w + 0.5 * hs * k1w, g_ex_mid, g_in_mid, I_ext, &k2V, &k2w);
            neuron_rhs(p, V + 0.5 * hs * k2V, w + 0.5 * hs * k2w, g_ex_mid, g_in_mid, I_ext, &k3V, &k3w);
            neuron_rhs(p, V + hs * k3V, w + hs * k3w, g_ex_next, g_in_next, I_ext, &k4V, &k4w);
            V += (hs / 6.0) * (k1V + 2.0 * k2V + 2.0 * k3V + k4V);
            w += (hs / 6.0) * (k1w + 2.0 * k2w + 2.0 * k3w + k4w);
        }
        g_ex = g_ex_next; dg_ex = dg_ex_next;
        g_in = g_in_next; dg_in = dg_in_next;
        if (refractory > 0) {
            --refractory;
        } else if (V >= p->V_peak) {
            V = V_reset;
            w += b;
            refractory = (int)(p->t_ref / hs + 0.5);
            spiked = 1;
        }
    }
    st[0] = V; st[1] = w; st[2] = g_ex; st[3] = dg_ex; st[4] = g_in; st[5] = dg_in;
    *refr = refractory;
    return spiked;
}

/* Tsodyks-Markram update of a presynaptic neuron on a spike at t_now
 * (src:565-576); returns the release multiplier x*u. */
static double stp_on_spike(double* x, double* u, double* t_last, double t_now, double U, double tau_rec,
                           double tau_fac) {
    double elapsed = t_now - *t_last;
    double x_decay = as_exp(-elapsed / tau_rec);
    double u_decay = tau_fac < 1e-10 ? 0.0 : as_exp(-elapsed / tau_fac);
    *x = 1.0 + (*x - *x * *u - 1.0) * x_decay;
    *u = U + *u * (1.0 - U) * u_decay;
    *t_last = t_now;
    return *x * *u;
}

/* neuron_step_kernel body for neuron j (src:433-587); returns 1 on a spike */
static int neuron_cell(orc_sim* s, uint32_t j, int64_t step, uint64_t noise_idx, int sic_refresh) {
    const orc_net* net = s->net;
    int excitatory = j < net->n_exc;
    const nrn_par* p = excitatory ? &s->pe : &s->pi;
    const csr_t* ce = &net->csr[0];
    const csr_t* ci = &net->csr[1];
    const csr_t* cs = &net->csr[2];
    int slot_ex = (int)((step + s->R - s->d_e) % s->R);
    const double* row_ex = s->ring_neuron + (size_t)slot_ex * net->n_neurons;
    double exc_sum = gather_sum(ce, row_ex, j);
    int slot_in = (int)((step + s->R - s->d_i) % s->R);
    const double* row_in = s->ring_neuron + (size_t)slot_in * net->n_neurons;
    /* the stored source id is the local inhibitory index, used as-is (src:461-462) */
    double inh_sum = gather_sum(ci, row_in, j);
    double exc_input = s->w_e * exc_sum;
    double inh_input = (-s->w_i) * inh_sum;
    double rate = excitatory ? 2700.0 : 2500.0;
    if (rate > 0.0) {
        double lambda = rate * s->dt * 1e-3;
        int ev = orc_rng_poisson(net->seed ^ 0x2545F4914F6CDD1DULL, (uint64_t)step * 1000003ULL + j, lambda);
        if (ev > 0) exc_input += 1.0 * (double)ev;
    }
    if (sic_refresh) {
        int slot_sic = (int)((step + s->R - s->d_a2n) % s->R);
        const double* row_sic = s->ring_sic + (size_t)slot_sic * net->n_astro;
        s->I_sic_hold[j] = s->w_a2n * gather_sum(cs, row_sic, j);
    }
    double I_ext = s->I_sic_hold[j];
    const double noise_std = 100.0;
    if (noise_std > 0.0) {
        I_ext += noise_std * orc_rng_normal(net->seed ^ 0x9E3779B185EBCA87ULL, noise_idx * 1000033ULL + (uint64_t)j);
    }
    double st[6] = {s->V[j], s->w[j], s->g_ex[j], s->dg_ex[j] + exc_input * p->psc_ex,
                    s->g_in[j], s->dg_in[j] + inh_input * p->psc_in};
    int32_t refractory = s->refr[j];
    int spiked = neuron_substeps(p, st, &refractory, I_ext, net->V_reset[j], net->b[j], s->h_step, s->substeps);
    s->V[j] = st[0]; s->w[j] = st[1];
    s->g_ex[j] = st[2]; s->dg_ex[j] = st[3];
    s->g_in[j] = st[4]; s->dg_in[j] = st[5];
    s->refr[j] = refractory;
    double out = 0.0;
    if (spiked) {
        if (s->stp_enabled) {
            const double U = 0.5, tau_rec = 800.0, tau_fac = 0.0; /* src:768 */
            out = stp_on_spike(&s->stp_x[j], &s->stp_u[j], &s->stp_t_last[j], (double)step * s->dt, U, tau_rec,
                               tau_fac);
        } else {
            out = 1.0;
        }
    }
    s->ring_neuron[(size_t)(step % s->R) * net->n_neurons + j] = out;
    return spiked;
}

void orc_sim_step(orc_sim* s, int64_t k, uint8_t* raster) {
    const orc_net* net = s->net;
    const int64_t na = net->n_astro, nn = net->n_neurons;
#ifdef _OPENMP
    int nt = s->threads > 0 ? s->threads : omp_get_max_threads();
#endif
    for (int64_t it = 0; it < k; ++it) {
        const int64_t step = s->step;
        /* host loop scalars (src:1043-1046) */
        uint64_t noise_idx = (uint64_t)(step / s->steps_per_noise);
        int sic_refresh = (step - s->d_a2n >= 0) && ((step - s->d_a2n) % 1 == 0);
        uint64_t spikes = 0;
        uint8_t* rrow = raster ? raster + (size_t)it * (size_t)nn : NULL;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
        {
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
            for (int64_t a = 0; a < na; ++a) astro_cell(s, (uint32_t)a, step);
#ifdef _OPENMP
#pragma omp for schedule(static) reduction(+ : spikes)
#endif
            for (int64_t j = 0; j < nn; ++j) {
                int sp = neuron_cell(s, (uint32_t)j, step, noise_idx, sic_refresh);
                spikes += (uint64_t)sp;
                if (rrow) rrow[j] = (uint8_t)sp;
            }
        }
        s->spikes += spikes;
        s->step++;
    }
}

int64_t orc_sim_steps_done(const orc_sim* s) { return s->step; }
uint64_t orc_sim_spike_count(const orc_sim* s) { return s->spikes; }
void orc_sim_reset_ever_active(orc_sim* s) { memset(s->ever_active, 0, s->net->n_astro); }

void orc_sim_read(const orc_sim* s, int field, double* out) {
    const double* f[13] = {s->Ca, s->IP3, s->h, s->V, s->w, s->g_ex, s->dg_ex, s->g_in,
                           s->dg_in, s->I_sic_hold, s->stp_x, s->stp_u, s->stp_t_last};
    size_t n = field < 3 ? s->net->n_astro : s->net->n_neurons;
    memcpy(out, f[field], n * sizeof(double));
}
void orc_sim_read_refractory(const orc_sim* s, int32_t* out) {
    memcpy(out, s->refr, s->net->n_neurons * sizeof(int32_t));
}
void orc_sim_read_ever_active(const orc_sim* s, uint8_t* out) {
    memcpy(out, s->ever_active, s->net->n_astro);
}

/* ======================================================================
 * Per-cell entry points for the known-answer tests of the step arithmetic
 * (the same functions the stepper above calls).
 * ====================================================================== */
void orc_astro_rhs(double Ca_tot, double IP3_0, double tau_IP3, double Ca, double IP3, double h, double noise,
                   double out[3]) {
    astro_rhs(Ca_tot, IP3_0, tau_IP3, Ca, IP3, h, noise, &out[0], &out[1], &out[2]);
}

void orc_astro_advance(double Ca_tot, double IP3_0, double tau_IP3, double delta_IP3, double ip3_input,
                       double noise, double h_step, int substeps, double state[3]) {
    astro_advance(Ca_tot, IP3_0, tau_IP3, delta_IP3, ip3_input, noise, h_step, substeps, &state[0], &state[1],
                  &state[2]);
}

double orc_sic_factor(double Ca) { return sic_factor(Ca); }

void orc_neuron_rhs(int excitatory, double V, double w, double g_ex, double g_in, double I_ext, double out[2]) {
    nrn_par p = excitatory ? exc_par() : inh_par();
    neuron_rhs(&p, V, w, g_ex, g_in, I_ext, &out[0], &out[1]);
}

int orc_neuron_substeps(int excitatory, double state[6], int32_t* refractory, double I_ext, double V_reset,
                        double b, double h_step, int substeps) {
    nrn_par p = excitatory ? exc_par() : inh_par();
    return neuron_substeps(&p, state, refractory, I_ext, V_reset, b, h_step, substeps);
}

double orc_psc_init(int excitatory, int inhibitory_synapse) {
    nrn_par p = excitatory ? exc_par() : inh_par();
    return inhibitory_synapse ? p.psc_in : p.psc_ex;
}

double orc_stp_on_spike(double state[3], double t_now) {
    return stp_on_spike(&state[0], &state[1], &state[2], t_now, 0.5, 800.0, 0.0);
}

double orc_gather(const uint32_t* row_start, const uint32_t* src, uint32_t n_targets, const double* ring_row,
                  uint32_t target) {
    csr_t c;
    c.row_start = (uint32_t*)row_start;
    c.src = (uint32_t*)src;
    c.n_targets = n_targets;
    c.n_edges = row_start[n_targets];
    return gather_sum(&c, ring_row, target);
}

void orc_sim_params(const orc_sim* s, int32_t out[5]) {
    out[0] = s->d_e; out[1] = s->d_i; out[2] = s->d_a2n; out[3] = s->R; out[4] = s->steps_per_noise;
}
