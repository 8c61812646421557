/* oracle_main.c -- command-line front end of the FP64 host oracle.
 *
 * TEST INFRASTRUCTURE / CPU BASELINE ONLY. Same flags and summary lines as
 * the reference program (src/astrosim_fused.cu:791-860, 1020-1143), with the
 * step loop running on host cores (OpenMP). The wall-clock window covers the
 * same steps as the reference's (src:1040, src:1104).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char** argv) {
    uint32_t n_total = 1000;
    uint64_t seed = 1;
    double dt = 0.1, pre_ms = 1000.0, sim_ms = 5000.0;
    int substeps = 1, threads = 0;
    double w_n2a = NAN, w_a2n = NAN;
    int stp = 1;
    for (int i = 1; i < argc; ++i) {
#define ARG(name) (strcmp(argv[i], name) == 0)
#define NEXT() ((i + 1 < argc) ? argv[++i] : "")
        if (ARG("--n") || ARG("-n")) n_total = (uint32_t)atoll(NEXT());
        else if (ARG("--seed")) seed = (uint64_t)atoll(NEXT());
        else if (ARG("--dt")) dt = atof(NEXT());
        else if (ARG("--substeps")) substeps = atoi(NEXT());
        else if (ARG("--pre-ms")) pre_ms = atof(NEXT());
        else if (ARG("--sim-ms")) sim_ms = atof(NEXT());
        else if (ARG("--block")) (void)NEXT(); /* accepted, no meaning on the host */
        else if (ARG("--threads")) threads = atoi(NEXT());
        else if (ARG("--w-n2a")) w_n2a = atof(NEXT());
        else if (ARG("--w-a2n")) w_a2n = atof(NEXT());
        else if (ARG("--no-stp")) stp = 0;
        else if (ARG("--help") || ARG("-h")) {
            printf("astrosim_oracle --n N_total [--seed S] [--dt 0.1] [--substeps 1] [--pre-ms 1000]\n"
                   "                [--sim-ms 5000] [--threads T] [--w-n2a W] [--w-a2n W] [--no-stp]\n"
                   "FP64 host oracle of the astrosim_fused step pipeline (test/baseline use).\n");
            return 0;
        } else {
            fprintf(stderr, "unknown argument: %s (try --help)\n", argv[i]);
            return 1;
        }
    }
    orc_net* net = orc_net_build(n_total, seed);
    if (!net) { fprintf(stderr, "invalid --n\n"); return 1; }
    uint32_t c[4];
    double p;
    orc_net_sizes(net, c, &p);
    printf("=== astrosim_oracle: FP64 host oracle of astrosim_fused ===\n");
    printf("N_total=%u  N_E=%u  N_I=%u  N_A=%u  p_primary=%.6g  (fixed in-degree law, K_SYN=%.0f)\n",
           n_total, c[0], c[1], c[2], p, 80.0);
    printf("w_n2a=%g  w_a2n=%g%s\n", isnan(w_n2a) ? 0.2 : w_n2a, isnan(w_a2n) ? 1.0 : w_a2n,
           (!isnan(w_n2a) || !isnan(w_a2n)) ? "  (overridden)" : "");
    printf("dt=%g ms  substeps=%d  pre=%g ms  sim=%g ms  seed=%llu  precision=%s\n", dt, substeps,
           pre_ms, sim_ms, (unsigned long long)seed, "double");
    printf("connectivity: exc=%llu inh=%llu e2a=%llu sic=%llu edges, built in %.3f s\n",
           (unsigned long long)orc_net_csr(net, 0, NULL, NULL), (unsigned long long)orc_net_csr(net, 1, NULL, NULL),
           (unsigned long long)orc_net_csr(net, 3, NULL, NULL), (unsigned long long)orc_net_csr(net, 2, NULL, NULL),
           orc_net_build_seconds(net));
    int nt = threads;
#ifdef _OPENMP
    if (nt <= 0) nt = omp_get_max_threads();
#else
    nt = 1;
#endif
    printf("CPU: host oracle, %d thread(s)\n", nt);
    orc_opts o = {dt, substeps, w_n2a, w_a2n, stp, nt};
    orc_sim* sim = orc_sim_create(net, &o);

    const int64_t pre_steps = (int64_t)(pre_ms / dt + 0.5);
    const int64_t sim_steps = (int64_t)(sim_ms / dt + 0.5);
    const int64_t total = pre_steps + sim_steps;
    const int64_t ckpt = sim_steps / 10 > 1 ? sim_steps / 10 : 1;
    int64_t print_every = total / 20 > 1 ? total / 20 : 1;
    uint64_t before = 0, at_last = 0;
    int64_t steps_last = pre_steps;
    double t0 = now_s();
    for (int64_t step = 0; step < total; ++step) {
        orc_sim_step(sim, 1, NULL);
        if (step == pre_steps - 1 || (pre_steps == 0 && step == 0)) {
            before = orc_sim_spike_count(sim);
            at_last = before;
            orc_sim_reset_ever_active(sim);
        }
        if (step >= pre_steps && (step - pre_steps) % ckpt == 0 && step > steps_last) {
            uint64_t now = orc_sim_spike_count(sim);
            double ws = (double)(step - steps_last) * dt * 1e-3;
            double hz = (double)(now - at_last) / ((double)c[3] * ws);
            printf("  step %8lld / %8lld (%5.1f%%)  windowed rate: %.4f Hz\n", (long long)step,
                   (long long)total, 100.0 * step / total, hz);
            fflush(stdout);
            at_last = now;
            steps_last = step;
        } else if (step % print_every == 0) {
            printf("  step %8lld / %8lld (%5.1f%%)\r", (long long)step, (long long)total, 100.0 * step / total);
            fflush(stdout);
        }
    }
    double wall = now_s() - t0;
    uint64_t measured = orc_sim_spike_count(sim) - before;
    double* ca = (double*)malloc(c[2] * sizeof(double));
    uint8_t* ever = (uint8_t*)malloc(c[2]);
    orc_sim_read(sim, 0, ca);
    orc_sim_read_ever_active(sim, ever);
    double mean_ca = 0.0;
    uint32_t active = 0, engaged = 0;
    for (uint32_t a = 0; a < c[2]; ++a) {
        mean_ca += ca[a];
        if ((ca[a] - 0.19669) * 1000.0 > 1.0) ++active;
        engaged += ever[a];
    }
    mean_ca /= (c[2] > 1 ? c[2] : 1);
    double rate = (double)measured / ((double)c[3] * (sim_ms * 1e-3));
    printf("\n=== done ===\n");
    printf("wall time            : %.3f s for %lld steps (%.2f us/step, both kernels)\n", wall,
           (long long)total, wall * 1e6 / (double)total);
    printf("spikes (recorded win): %llu\n", (unsigned long long)measured);
    printf("mean firing rate     : %.4f Hz\n", rate);
    printf("mean calcium (final) : %.6g uM  (SIC_th=%.6g uM)\n", mean_ca, 0.19669);
    printf("astrocytes active    : %u / %u (Ca above SIC threshold at the final step only)\n", active, c[2]);
    printf("astrocytes engaged   : %u / %u (Ca crossed SIC threshold at ANY point in the recorded window -- "
           "comparable to astrosimgpu's 'astrocytes with transients')\n", engaged, c[2]);
    printf("steps/s              : %.3f\n", (double)total / wall);
    free(ca); free(ever);
    orc_sim_free(sim);
    orc_net_free(net);
    return 0;
}
