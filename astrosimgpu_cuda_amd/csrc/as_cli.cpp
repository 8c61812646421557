// as_cli.cpp -- `astrosim_fused`, the drop-in command line.
//
// Same flags, defaults, banner, progress and summary lines as the reference
// program (src/astrosim_fused.cu:791-1143), driving the GPU through the C
// ABI. Extra flags select this build's precision and method; extra summary
// lines come after the reference's so line-oriented parsers keep working.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/astrosim.h"

namespace {

int fail(const char* what) {
    std::fprintf(stderr, "%s: %s\n", what, as_last_error());
    std::abort();  // reference convention: runtime API errors abort (src:109-117)
}

}  // namespace

int main(int argc, char** argv) {
    uint32_t n_total = 1000;
    uint64_t seed = 1;
    double dt = 0.1;
    int substeps = 1;
    double pre_ms = 1000.0, sim_ms = 5000.0;
    int block = 128;
    double w_n2a = std::nan(""), w_a2n = std::nan("");
    bool no_stp = false;
    std::string dump_edges;
    // blocked by default: bit-identical to the plain method (the reference's
    // per-step data flow), and several times faster
    int precision = AS_FP64, method = AS_METHOD_BLOCKED;
    int setup_where = -1;  // default: GPU generator, host walk for --dump-edges

    for (int i = 1; i < argc; ++i) {
        auto arg = [&](const char* name) { return std::strcmp(argv[i], name) == 0; };
        auto next = [&]() -> const char* { return (i + 1 < argc) ? argv[++i] : ""; };
        if (arg("--n") || arg("-n")) n_total = static_cast<uint32_t>(std::atoll(next()));
        else if (arg("--seed")) seed = static_cast<uint64_t>(std::atoll(next()));
        else if (arg("--dt")) dt = std::atof(next());
        else if (arg("--substeps")) substeps = std::atoi(next());
        else if (arg("--pre-ms")) pre_ms = std::atof(next());
        else if (arg("--sim-ms")) sim_ms = std::atof(next());
        else if (arg("--block")) block = std::atoi(next());
        else if (arg("--w-n2a")) w_n2a = std::atof(next());
        else if (arg("--w-a2n")) w_a2n = std::atof(next());
        else if (arg("--no-stp")) no_stp = true;
        else if (arg("--dump-edges")) dump_edges = next();
        else if (arg("--precision")) {
            const std::string v = next();
            if (v == "fp64") precision = AS_FP64;
            else if (v == "fp32c") precision = AS_FP32C;
            else { std::fprintf(stderr, "unknown --precision %s (fp64|fp32c)\n", v.c_str()); return 1; }
        } else if (arg("--method")) {
            const std::string v = next();
            if (v == "plain") method = AS_METHOD_PLAIN;
            else if (v == "blocked") method = AS_METHOD_BLOCKED;
            else { std::fprintf(stderr, "unknown --method %s (plain|blocked)\n", v.c_str()); return 1; }
        } else if (arg("--setup")) {
            const std::string v = next();
            if (v == "host") setup_where = AS_BUILD_HOST;
            else if (v == "gpu") setup_where = AS_BUILD_GPU;
            else { std::fprintf(stderr, "unknown --setup %s (host|gpu)\n", v.c_str()); return 1; }
        } else if (arg("--help") || arg("-h")) {
            std::printf(
                "astrosim_fused --n N_total [--seed S] [--dt 0.1] [--substeps 1] "
                "[--pre-ms 1000] [--sim-ms 5000] [--block 128]\n"
                "         [--w-n2a W] [--w-a2n W] [--no-stp] [--dump-edges PATH]\n"
                "         [--precision fp64|fp32c] [--method plain|blocked] [--setup host|gpu]\n"
                "N_total is split by the fixed-in-degree scaling law: N_E=0.8*N, N_I=0.2*N,\n"
                "N_A=N/5, p_primary=80/N_E. --w-n2a/--w-a2n override the neuron->astrocyte /\n"
                "astrocyte->neuron coupling weight (default 0.2 / 1.0); 0 disconnects one\n"
                "direction. --no-stp disables Tsodyks-Markram STP (multiplier 1.0).\n"
                "--dump-edges PATH writes the connectivity as CSV (type,source,target) and\n"
                "exits (on the host unless --setup gpu is given). --block is the workgroup size (multiple of 64).\n"
                "--precision fp32c runs FP32 state with compensated Ca/IP3/h. --method blocked\n"
                "(the default) advances one minimum-delay block of steps per launch; --method\n"
                "plain visits every in-edge every step, as the reference; both give the same\n"
                "FP64 results bit for bit (MI355X gfx950).\n"
                "--setup gpu (the default except with --dump-edges) generates the connectivity\n"
                "on the GPU, identical to --setup host's sequential walk.\n");
            return 0;
        } else {
            std::fprintf(stderr, "unknown argument: %s (try --help)\n", argv[i]);
            return 1;
        }
    }

    if (setup_where < 0) setup_where = dump_edges.empty() ? AS_BUILD_GPU : AS_BUILD_HOST;
    as_network* net = nullptr;
    const auto t_conn0 = std::chrono::steady_clock::now();
    if (as_network_build_on(n_total, seed, setup_where, &net) != AS_OK) {
        std::fprintf(stderr, "network: %s\n", as_last_error());
        return 1;
    }
    const auto t_conn1 = std::chrono::steady_clock::now();
    as_network_info info{};
    as_network_info_get(net, &info);
    const double used_w_n2a = std::isnan(w_n2a) ? 0.2 : w_n2a;
    const double used_w_a2n = std::isnan(w_a2n) ? 1.0 : w_a2n;

    std::printf("=== astrosim_fused: race-free, kernel-fused HIP astrosimgpu (MI355X) ===\n");
    std::printf("N_total=%u  N_E=%u  N_I=%u  N_A=%u  p_primary=%.6g  (fixed in-degree law, K_SYN=%.0f)\n",
                n_total, info.n_exc, info.n_inh, info.n_astro, info.p_primary, 80.0);
    std::printf("w_n2a=%g  w_a2n=%g%s\n", used_w_n2a, used_w_a2n,
                (!std::isnan(w_n2a) || !std::isnan(w_a2n)) ? "  (overridden)" : "");
    std::printf("dt=%g ms  substeps=%d  pre=%g ms  sim=%g ms  seed=%llu  precision=%s\n", dt, substeps,
                pre_ms, sim_ms, static_cast<unsigned long long>(seed),
                precision == AS_FP64 ? "double" : "float-compensated");
    (void)t_conn0;
    (void)t_conn1;
    std::printf("connectivity: exc=%llu inh=%llu e2a=%llu sic=%llu edges, built in %.3f s\n",
                (unsigned long long)info.n_exc_edges, (unsigned long long)info.n_inh_edges,
                (unsigned long long)info.n_e2a_edges, (unsigned long long)info.n_sic_edges,
                info.build_seconds);

    if (!dump_edges.empty()) {
        if (as_network_dump_edges(net, dump_edges.c_str()) != AS_OK) {
            std::fprintf(stderr, "cannot open %s for --dump-edges\n", dump_edges.c_str());
            return 1;
        }
        std::printf("Wrote edge list to %s\n", dump_edges.c_str());
        return 0;
    }

    char name[256], arch[64];
    const as_status ds = as_device_info(name, sizeof name, arch, sizeof arch);
    if (ds == AS_ERR_NODEVICE) {
        std::fprintf(stderr, "No GPU device visible. Run on a node with an MI355X (gfx950).\n");
        return 1;
    }
    if (ds != AS_OK) fail("device query");
    std::printf("GPU: %s (%s)\n", name, arch);

    as_config cfg;
    as_config_default(&cfg);
    cfg.dt = dt;
    cfg.substeps = substeps;
    cfg.block = block;
    cfg.w_n2a = w_n2a;
    cfg.w_a2n = w_a2n;
    cfg.stp_enabled = no_stp ? 0 : 1;
    cfg.precision = precision;
    cfg.method = method;
    as_sim* sim = nullptr;
    const as_status cs = as_sim_create(net, &cfg, &sim);
    if (cs == AS_ERR_INVALID) {
        std::fprintf(stderr, "invalid configuration: %s\n", as_last_error());
        return 1;
    }
    if (cs != AS_OK) fail("simulation setup");
    as_network_free(net);

    // ---- time loop (src/astrosim_fused.cu:1020-1102) ----
    const int64_t pre_steps = static_cast<int64_t>(pre_ms / dt + 0.5);
    const int64_t sim_steps = static_cast<int64_t>(sim_ms / dt + 0.5);
    const int64_t total_steps = pre_steps + sim_steps;
    const int64_t print_every = std::max<int64_t>(1, total_steps / 20);
    const int64_t checkpoint_every = std::max<int64_t>(1, sim_steps / 10);
    const uint32_t n_neurons = info.n_neurons;
    uint64_t spikes_before = 0, spikes_at_ckpt = 0;
    int64_t steps_at_ckpt = pre_steps;
    if (as_sim_sync(sim) != AS_OK) fail("sync");
    const auto t_run0 = std::chrono::steady_clock::now();
    int64_t enqueued = 0;
    auto run_through = [&](int64_t step) {  // enqueue steps up to and including `step`
        if (step + 1 > enqueued) {
            if (as_sim_step(sim, step + 1 - enqueued) != AS_OK) fail("step");
            enqueued = step + 1;
        }
    };
    for (int64_t step = 0; step < total_steps; ++step) {
        if (step == pre_steps - 1 || (pre_steps == 0 && step == 0)) {
            run_through(step);
            if (as_sim_spike_count(sim, &spikes_before) != AS_OK) fail("spike count");
            spikes_at_ckpt = spikes_before;
            if (as_sim_reset_ever_active(sim) != AS_OK) fail("reset");
        }
        if (step >= pre_steps && (step - pre_steps) % checkpoint_every == 0 && step > steps_at_ckpt) {
            run_through(step);
            uint64_t now = 0;
            if (as_sim_spike_count(sim, &now) != AS_OK) fail("spike count");
            const int64_t window_steps = step - steps_at_ckpt;
            const double window_s = window_steps * dt * 1e-3;
            const double hz = static_cast<double>(now - spikes_at_ckpt) / (static_cast<double>(n_neurons) * window_s);
            std::printf("  step %8lld / %8lld (%5.1f%%)  windowed rate: %.4f Hz\n", static_cast<long long>(step),
                        static_cast<long long>(total_steps), 100.0 * step / total_steps, hz);
            std::fflush(stdout);
            spikes_at_ckpt = now;
            steps_at_ckpt = step;
        } else if (step % print_every == 0) {
            run_through(step);
            std::printf("  step %8lld / %8lld (%5.1f%%)\r", static_cast<long long>(step),
                        static_cast<long long>(total_steps), 100.0 * step / total_steps);
            std::fflush(stdout);
        }
    }
    if (total_steps > enqueued) run_through(total_steps - 1);
    if (as_sim_sync(sim) != AS_OK) fail("sync");
    const auto t_run1 = std::chrono::steady_clock::now();

    uint64_t spikes_total = 0;
    if (as_sim_spike_count(sim, &spikes_total) != AS_OK) fail("spike count");
    const uint64_t spikes_measured = spikes_total - spikes_before;
    std::vector<double> ca(info.n_astro);
    std::vector<uint8_t> ever(info.n_astro);
    if (as_sim_read_state(sim, 0, ca.data()) != AS_OK) fail("read Ca");
    if (as_sim_read_ever_active(sim, ever.data()) != AS_OK) fail("read ever_active");
    const double SIC_th = 0.19669;
    double mean_ca = 0.0;
    uint32_t active = 0, engaged = 0;
    for (double c : ca) {
        mean_ca += c;
        if ((c - SIC_th) * 1000.0 > 1.0) ++active;
    }
    mean_ca /= std::max<uint32_t>(1, info.n_astro);
    for (uint8_t v : ever) engaged += v;
    const double wall_s = std::chrono::duration<double>(t_run1 - t_run0).count();
    const double rate = static_cast<double>(spikes_measured) / (static_cast<double>(n_neurons) * (sim_ms * 1e-3));

    std::printf("\n=== done ===\n");
    std::printf("wall time            : %.3f s for %lld steps (%.2f us/step, both kernels)\n", wall_s,
                static_cast<long long>(total_steps), wall_s * 1e6 / static_cast<double>(total_steps));
    std::printf("spikes (recorded win): %llu\n", static_cast<unsigned long long>(spikes_measured));
    std::printf("mean firing rate     : %.4f Hz\n", rate);
    std::printf("mean calcium (final) : %.6g uM  (SIC_th=%.6g uM)\n", mean_ca, SIC_th);
    std::printf("astrocytes active    : %u / %u (Ca above SIC threshold at the final step only)\n", active,
                info.n_astro);
    std::printf("astrocytes engaged   : %u / %u (Ca crossed SIC threshold at ANY point in the "
                "recorded window -- comparable to astrosimgpu's 'astrocytes with transients')\n",
                engaged, info.n_astro);
    std::printf("steps/s              : %.3f  (method=%s, device bytes=%llu)\n",
                static_cast<double>(total_steps) / wall_s, method == AS_METHOD_PLAIN ? "plain" : "blocked",
                static_cast<unsigned long long>(as_sim_bytes(sim)));
    as_sim_destroy(sim);
    return 0;
}
