/* astrosim.h -- C ABI of the MI355X-native astrosim_fused step pipeline.
 *
 * The reference program (src/astrosim_fused.cu) has no library API: its only
 * interface is the process command line and the terminal summary. Its main()
 * is split here into the calls a host program (this project's own
 * `astrosim_fused` CLI, the ensemble runner, the server, or any FFI caller)
 * needs, each citing the reference span it replaces. Plain C types and POD
 * pointers only; no C++ types cross the boundary and nothing throws. Every
 * call returns an as_status; as_last_error() gives the message of the last
 * failure on the calling thread.
 *
 * Threading: one as_sim is driven by one host thread at a time. Different
 * as_sim objects may be driven concurrently from different threads; each
 * owns its own HIP stream.
 */
#ifndef ASTROSIM_H
#define ASTROSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum as_status {
    AS_OK = 0,
    AS_ERR_INVALID = 1,  /* bad argument or unsupported configuration */
    AS_ERR_HIP = 2,      /* HIP runtime failure (message names file:line) */
    AS_ERR_NOMEM = 3,    /* device or host allocation failed */
    AS_ERR_NODEVICE = 4, /* no GPU visible (reference: src/astrosim_fused.cu:904-912) */
    AS_ERR_IO = 5,       /* file could not be written */
    AS_ERR_RANGE = 6     /* requested data is not (or no longer) held */
} as_status;

typedef enum as_precision {
    AS_FP64 = 0,  /* IEEE double throughout; bit-identical to the FP64 host oracle */
    AS_FP32C = 1  /* FP32 state and RHS, compensated hi+lo Ca/IP3/h, FP64 AdEx exponential */
} as_precision;

typedef enum as_method {
    AS_METHOD_PLAIN = 0,  /* one launch per step, dense delay rings (reference data flow) */
    AS_METHOD_BLOCKED = 1 /* one launch per min-delay block, bit-packed spike history */
} as_method;

typedef struct as_network as_network;
typedef struct as_sim as_sim;

/* Run options, the reference's command-line flags (src/astrosim_fused.cu:792-851). */
typedef struct as_config {
    double dt;          /* --dt, ms (default 0.1) */
    int32_t substeps;   /* --substeps (default 1) */
    int32_t block;      /* --block, threads per workgroup: 64..1024, multiple of 64 (default 128) */
    double w_n2a;       /* --w-n2a; NaN keeps the default 0.2 */
    double w_a2n;       /* --w-a2n; NaN keeps the default 1.0 */
    int32_t stp_enabled;/* 0 for --no-stp */
    int32_t precision;  /* as_precision */
    int32_t method;     /* as_method */
} as_config;

/* Fills the defaults above (FP64, plain method). */
void as_config_default(as_config* cfg);

/* Sizes and edge counts of a built network. */
typedef struct as_network_info {
    uint32_t n_total, n_exc, n_inh, n_astro, n_neurons;
    double p_primary;
    uint64_t n_exc_edges, n_inh_edges, n_e2a_edges, n_sic_edges;
    double build_seconds; /* connectivity generation + inversion, as printed at src:869-871 */
    int32_t built_on_gpu; /* 1 when as_network_build_on generated the connectivity on the GPU */
} as_network_info;

/* Where as_network_build_on generates the connectivity. */
typedef enum as_build_where {
    AS_BUILD_HOST = 0,     /* sequential host walk (as_network_build) */
    AS_BUILD_GPU = 1,      /* parallel generator on the current GPU from N_total = 100000 on (below,
                              and where it declines, the host walk, which is then as fast) */
    AS_BUILD_GPU_ONLY = 2  /* parallel generator only: AS_ERR_INVALID when it declines */
} as_build_where;

/* ---- network construction (as_network_build: host only, no GPU needed) ---- */

/* Scaling law, connectivity and per-cell parameters for (n_total, seed).
 * Replaces src/astrosim_fused.cu:846 (ScalingConfig::make), :867
 * (build_connectivity) and :916-933 (population draws). */
as_status as_network_build(uint32_t n_total, uint64_t seed, as_network** out);
/* The same network, bit for bit, with the connectivity generated where
 * `where` (as_build_where) says. The GPU generator walks the reference's one
 * sequential setup stream in parallel (speculative chains stitched where
 * they meet, autapse repairs, an exact walk of every row that proves the
 * result) and inverts the edges with stable radix sorts; it declines
 * networks with N_total below a few thousand. Same spans as as_network_build. */
as_status as_network_build_on(uint32_t n_total, uint64_t seed, int32_t where, as_network** out);
as_status as_network_info_get(const as_network* net, as_network_info* info);
/* CSR by target. which: 0 exc, 1 inh, 2 sic (astro->neuron), 3 e2a (neuron->astro).
 * Either pointer may be NULL; row_start needs n_targets+1 entries.
 * Replaces HostCSR (src/astrosim_fused.cu:304-315). */
as_status as_network_csr(const as_network* net, int which, uint32_t* row_start, uint32_t* src);
/* Per-cell parameters. which: 0 Ca_tot 1 IP3_0 2 tau_IP3 3 delta_IP3 (n_astro values),
 * 4 V_reset 5 b (n_neurons values). Replaces src/astrosim_fused.cu:918-933. */
as_status as_network_params(const as_network* net, int which, double* out);
/* --dump-edges CSV writer. Replaces src/astrosim_fused.cu:873-902. */
as_status as_network_dump_edges(const as_network* net, const char* path);
void as_network_free(as_network* net);

/* ---- simulation on the GPU ---- */

/* Device memory a simulation of this network would take, in bytes
 * (CSR + state + rings), for admission control. */
as_status as_sim_footprint(const as_network* net, const as_config* cfg, uint64_t* bytes);
/* Allocates and uploads connectivity, parameters and initial state; the
 * network may be freed afterwards. Replaces src/astrosim_fused.cu:904-1018. */
as_status as_sim_create(const as_network* net, const as_config* cfg, as_sim** out);
/* Enqueues k steps on the simulation's stream and returns (asynchronous).
 * Replaces one pass of the time loop body, src/astrosim_fused.cu:1042-1064. */
as_status as_sim_step(as_sim* sim, int64_t k);
/* Waits for all enqueued work (src/astrosim_fused.cu:1103). */
as_status as_sim_sync(as_sim* sim);
/* Steps completed or enqueued so far. */
int64_t as_sim_steps(const as_sim* sim);
/* Total spikes since creation (synchronises). Replaces the spike-counter
 * reads at src/astrosim_fused.cu:1068, 1084, 1108. */
as_status as_sim_spike_count(as_sim* sim, uint64_t* out);
/* Clears every astrocyte's ever-active flag (enqueued). src/astrosim_fused.cu:1075. */
as_status as_sim_reset_ever_active(as_sim* sim);
/* Reads a state array (synchronises). field: 0 Ca 1 IP3 2 h (n_astro);
 * 3 V 4 w 5 g_ex 6 dg_ex 7 g_in 8 dg_in 9 I_sic_hold 10 stp_x 11 stp_u
 * 12 stp_t_last (n_neurons). Values are converted to double. Replaces the
 * final D2H reads at src/astrosim_fused.cu:1111-1112. */
as_status as_sim_read_state(as_sim* sim, int field, double* out);
as_status as_sim_read_refractory(as_sim* sim, int32_t* out);
/* Per-astrocyte ever-active flags (src/astrosim_fused.cu:1121-1123). */
as_status as_sim_read_ever_active(as_sim* sim, uint8_t* out);
/* Spikes of every neuron since creation (n_neurons values; synchronises).
 * The reference keeps only the total (src/astrosim_fused.cu:584); per-neuron
 * counts give the recorded window's synaptic events (each spike times its
 * neuron's out-degree) and per-neuron rates. */
as_status as_sim_read_spike_counts(as_sim* sim, uint32_t* out);
/* Spike raster of steps [first_step, first_step + n_steps): n_steps*n_neurons
 * bytes, 1 = spiked. Only the most recent steps are held (at least the
 * minimum synaptic delay); AS_ERR_RANGE otherwise. Test and analysis aid,
 * the reference has no counterpart. */
as_status as_sim_read_raster(as_sim* sim, int64_t first_step, int32_t n_steps, uint8_t* out);
/* Spike events of steps [first_step, first_step + n_steps) -- steps the
 * simulation still holds, as for as_sim_read_raster (AS_ERR_RANGE otherwise)
 * -- in step order and, within a step, neuron order: the first
 * min(total, cap) events go to steps[i] / neurons[i], and *count receives
 * the total (a caller that gets *count > cap calls again with more room).
 * The events are compacted on the device, in staging this call allocates and
 * frees (hipMalloc); synchronises. The reference keeps only a spike total
 * (src/astrosim_fused.cu:584); this is the recorder's read. */
as_status as_sim_read_spikes(as_sim* sim, int64_t first_step, int32_t n_steps, int64_t* steps, uint32_t* neurons,
                             uint64_t cap, uint64_t* count);
/* The same, allocating nothing: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under graph capture,
 * tight serving or recording loops). The events are staged in a fixed device
 * array of 2^20 events (8 MB, one per GPU, held until the process ends), so
 * cap may not exceed 1,048,576 (AS_ERR_INVALID otherwise); n_steps has no
 * limit. Every call on a GPU shares that GPU's array: calls that overlap on
 * one GPU (from several threads) race. */
as_status as_sim_read_spikes_noalloc(as_sim* sim, int64_t first_step, int32_t n_steps, int64_t* steps,
                                     uint32_t* neurons, uint64_t cap, uint64_t* count);
/* Device bytes held by this simulation. */
uint64_t as_sim_bytes(const as_sim* sim);
/* Kernel timing: when enabled, every launch of the method's dominant step
 * kernel (plain: the fused step kernel, one launch per step; blocked: the
 * integrate kernel, one launch per block or partial block) is bracketed by HIP events on the simulation's stream.
 * as_sim_kernel_time returns the summed milliseconds and launch count of the
 * timed launches since the previous read (synchronises). */
as_status as_sim_set_timing(as_sim* sim, int enable);
as_status as_sim_kernel_time(as_sim* sim, double* total_ms, int64_t* launches);
void as_sim_destroy(as_sim* sim);

/* ---- batches: many networks advanced together (blocked method) ----
 *
 * No counterpart in the reference, which runs one network per process. A
 * batch holds an HBM arena of fixed capacity; member networks occupy ranges
 * of it, join and leave at block boundaries (1 ms at dt 0.1), and advance
 * together in one integrate launch per block. A member's
 * results are bit-identical to running it alone (its random streams use its
 * own seed, local cell ids and local steps). Members are numbered by slot. */
typedef struct as_batch as_batch;
typedef struct as_batch_capacity {
    uint64_t neurons, astros;
    uint64_t edges[4]; /* exc, inh, sic, e2a */
    uint32_t members;
} as_batch_capacity;
typedef struct as_batch_info {
    int64_t global_step;      /* steps enqueued since creation */
    int32_t block_steps;      /* block length L in steps */
    int32_t active_members;
    int32_t block_threads;    /* workgroup (tile) size in use */
    uint64_t device_bytes;
    uint64_t free_neurons;    /* largest free neuron range */
    uint64_t bucket_overflows;/* targets that fell back to a full in-row walk */
    int32_t chunk_steps;      /* steps one integrate launch advances (LDS chunk) */
} as_batch_info;

/* Exact capacity for the given networks side by side. */
as_status as_batch_capacity_for(const as_network* const* nets, int32_t n, as_batch_capacity* cap);
as_status as_batch_footprint(const as_batch_capacity* cap, const as_config* cfg, uint64_t* bytes);
/* cfg->method must be AS_METHOD_BLOCKED. */
as_status as_batch_create(const as_batch_capacity* cap, const as_config* cfg, as_batch** out);
/* *fits = 1 when the network fits the free ranges and a member slot is free. */
as_status as_batch_fits(as_batch* b, const as_network* net, int32_t* fits);
/* Admits a network to run n_steps steps from the current global step (a
 * block boundary). Uploads it and builds its out-CSR on the batch's stream. */
as_status as_batch_add(as_batch* b, const as_network* net, int64_t n_steps, int32_t* member);
/* as_batch_add for n networks: the uploads run back to back and are waited
 * for once. If one does not fit, the networks before it are admitted, the
 * rest get member -1, and the error is returned. No counterpart in the
 * reference (it runs one network per process). */
as_status as_batch_add_many(as_batch* b, const as_network* const* nets, int32_t n, const int64_t* n_steps,
                            int32_t* members);
/* Retires a member; its ranges are reused once in-flight deliveries drain. */
as_status as_batch_remove(as_batch* b, int32_t member);
/* Advances all active members k steps (asynchronous; partial blocks allowed). */
as_status as_batch_step(as_batch* b, int64_t k);
/* Advances whole blocks by replaying a captured hipGraph (asynchronous). */
as_status as_batch_run_graph(as_batch* b, int64_t blocks);
/* Times each candidate workgroup size, each with the longest LDS chunk its
 * budget allows and with about half of it, over `blocks_each` blocks of the
 * live run (the results depend on neither), and keeps the fastest pair;
 * *chosen is the workgroup size (as_batch_info.chunk_steps the chunk). The
 * block length is the delay in steps (the longest the delay allows: every
 * emission is delivered by the next block). */
as_status as_batch_autotune(as_batch* b, const int32_t* candidates, int32_t n, int32_t blocks_each,
                            int32_t* chosen);
as_status as_batch_sync(as_batch* b);
as_status as_batch_get_info(as_batch* b, as_batch_info* info);
/* Steps the member has run and the steps it was admitted for. */
as_status as_batch_member_steps(as_batch* b, int32_t member, int64_t* done, int64_t* target);
as_status as_batch_spike_count(as_batch* b, int32_t member, uint64_t* out);
as_status as_batch_reset_ever_active(as_batch* b, int32_t member);
as_status as_batch_read_state(as_batch* b, int32_t member, int32_t field, double* out);
as_status as_batch_read_ever_active(as_batch* b, int32_t member, uint8_t* out);
as_status as_batch_read_raster(as_batch* b, int32_t member, int64_t first_step, int32_t n_steps, uint8_t* out);
/* Spikes of every neuron of the member since its admission. */
as_status as_batch_read_spike_counts(as_batch* b, int32_t member, uint32_t* out);
as_status as_batch_set_timing(as_batch* b, int enable);
as_status as_batch_kernel_time(as_batch* b, double* total_ms, int64_t* launches);
void as_batch_destroy(as_batch* b);

/* ---- continuous-batching server ----
 *
 * No counterpart in the reference. Requests are built by worker threads (by
 * default with the GPU generator from N_total = 100000 on, see
 * as_network_build_on), admitted into resident FP32c / FP64 batches
 * (capacities sized from a fraction of free HBM) as soon as their footprint
 * fits, advanced in rounds of blocks, and retired when they reach their
 * step count. Results equal a solo run. The server creates no more HIP
 * streams than the process has hardware queues (GPU_MAX_HW_QUEUES, HIP's
 * default 4; streams that share a queue wait for each other's work): a
 * block stream per batch from 5 queues on (else one for both), one upload
 * stream, and the rest for network builds, one builder per stream; a
 * builder finding none free builds networks below 400,000 cells on the
 * host. */
typedef struct as_server as_server;
typedef struct as_server_config {
    double dt;                /* ms */
    int32_t block;            /* workgroup size */
    double hbm_fraction;      /* share of free HBM given to the two batches */
    double fp64_share;        /* share of that budget given to the FP64 batch */
    int32_t build_threads;    /* host network builders */
    int32_t blocks_per_round; /* blocks per round (default 16); rounds are
                                 pipelined two deep, members join and leave
                                 between rounds */
    int32_t max_members;      /* member slots per batch */
    int32_t setup_where;      /* as_build_where of the request networks (default AS_BUILD_GPU) */
    int32_t fail_after_rounds;/* test hook: the driver stops with an error after this many
                                 rounds (0, the default: never) */
} as_server_config;
typedef struct as_request {
    uint32_t n_total;   /* --n */
    uint64_t seed;      /* --seed */
    int64_t n_steps;    /* steps to simulate */
    int32_t precision;  /* AS_FP32C or AS_FP64 */
} as_request;
enum { AS_REQ_BUILDING = 0, AS_REQ_QUEUED = 1, AS_REQ_RUNNING = 2, AS_REQ_DONE = 3, AS_REQ_REJECTED = 4 };
typedef struct as_request_result {
    int64_t id;
    int32_t status;               /* AS_REQ_* */
    uint32_t n_neurons, n_astro;
    double t_submit, t_built, t_admit, t_done;  /* seconds since the server started */
    uint64_t spikes;              /* spikes over the whole run */
    double mean_ca;               /* final mean calcium */
    uint32_t active_astro;        /* astrocytes above the SIC threshold at the end */
    double t_build_start;         /* when a builder took the request (seconds since the server started) */
    int32_t built_on_gpu;         /* 1: the GPU connectivity generator built the network; 0: the host walk */
} as_request_result;
typedef struct as_server_stats {
    uint64_t submitted, completed, rejected;
    double p50_latency_s, p99_latency_s, mean_latency_s;  /* submit -> done */
    double sims_per_s;            /* completed / (last done - first submit) */
    uint64_t device_bytes;        /* HBM held by the resident batches */
    uint64_t peak_active;         /* most requests resident at once */
    uint64_t peak_resident_cells, capacity_cells;
    double gpu_busy_s;            /* device time of the rounds (per round, the
                                     longer batch's span) */
    double cell_steps;            /* cells x steps of the completed requests */
    double admit_s, retire_s;     /* driver time admitting and retiring requests between rounds */
    int32_t dedicated_queues;     /* 1: every server stream has a hardware queue of its own */
    int32_t build_streams;        /* streams the network generator's builds run on */
} as_server_stats;

void as_server_config_default(as_server_config* cfg);
as_status as_server_create(const as_server_config* cfg, as_server** out);
as_status as_server_submit(as_server* s, const as_request* req, int64_t* id);
as_status as_server_result(as_server* s, int64_t id, as_request_result* out);
/* Waits until every submitted request is done or rejected (*all_done = 1) or
 * the timeout passes (*all_done = 0). If the driver stopped after an error
 * (a HIP failure), every unfinished request is rejected, the wait ends at
 * once and the error is returned (AS_ERR_HIP, message in as_last_error);
 * as_server_submit and as_server_get_stats return it too from then on. */
as_status as_server_wait_all(as_server* s, double timeout_s, int32_t* all_done);
as_status as_server_get_stats(as_server* s, as_server_stats* out);
void as_server_destroy(as_server* s);

/* ---- misc ---- */
/* Name of device 0 and its GCN arch ("gfx950"); AS_ERR_NODEVICE without a GPU. */
as_status as_device_info(char* name, int name_len, char* arch, int arch_len);
/* Free and total device memory in bytes. */
as_status as_device_memory(uint64_t* free_bytes, uint64_t* total_bytes);
const char* as_last_error(void);
/* Library version string. */
const char* as_version(void);

#ifdef __cplusplus
}
#endif

#endif /* ASTROSIM_H */
