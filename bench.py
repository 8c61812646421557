"""Headline benchmark: simulation steps per second of the astrosim_fused step
pipeline on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n N] [--seed S]
                    [--legs fp64:plain,...] [--no-cpu-baseline] [--no-ensemble]
                    [--serving-seconds S] [--no-config1] [--no-activity]

Workload (default): the full-scale single network, N_total = 10,000,000
(8M excitatory + 2M inhibitory neurons, 2M astrocytes, ~1.32G int32 in-edges),
seed 42, dt 0.1 ms. A "step" is one simulated 0.1 ms step of every cell. The
network is generated (on the GPU, identical to the host walk of the setup
stream) and resident in HBM before timing starts. Each leg (precision and
method) runs W untimed steps, then exactly K timed steps bracketed by device
syncs. The default K = 11,000 is the reference's run at this size (pre 100
ms + sim 1000 ms; its steps/s covers both phases), onset transient included.

One JSON line is printed. `value` is the headline leg's (the last leg's:
FP32c, blocked method) steps/s; every leg's figures are in `legs`:
  - steps/s over the timed window, and synaptic events/s: every spike of the
    window times its neuron's out-degree over the connections that deliver it
    (excitatory, inhibitory and neuron-to-astrocyte in-CSRs, counted by
    source id), over the window's wall time;
  - the dominant kernel's time per FULL launch (one step for the plain
    method, one 10-step block for the blocked method): its time summed over
    the timed window (HIP events around every launch on the simulation's
    stream) per 10 (blocked) or 1 (plain) steps, so that partial launches at
    the window's ends do not change the basis; and, separately, the average
    of 20 (blocked) or 100 (plain) whole launches shortly after the window
    (after 50 or 200 untimed ones);
  - that kernel's own algorithmic HBM bytes per launch (see
    kernel_bytes_per_launch), the GB/s they give at that launch time and the
    fraction of the 8 TB/s peak, beside the rocprofv3-measured HBM bytes
    (profiles/traffic_*.json: FETCH_SIZE and WRITE_SIZE scaled by a
    known-size copy) and, for the blocked kernels, the vector-ALU busy
    fraction (profiles/valu_*.json, SQ counters).
`roofline` repeats the headline leg's figures. `fp32c_vs_fp64` runs both
blocked legs over the reference's whole run (1,000 + 10,000 steps, untimed)
and compares the recorded window's firing rate and SIC fractions; the FP64
blocked path is bit-identical to the FP64 host oracle. `cpu_baseline` times
the FP64 host oracle's stepper on the same network. `config1` times the
reference's default-size run (N_total = 500, 60,000 steps), `activity` the
two FP64 methods on a 1M-cell network at normal and at SIC-driven high
activity, `ensemble` the 256-network batch and `serving` the 40 req/s load.

    python bench.py --copy-calibration BYTES   # the known-size copy alone
"""
from __future__ import annotations

import argparse
import json
import re
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# Blocked-method tile (threads per workgroup) measured fastest at N=10M: FP32c
# 128 (0.656 ms per launch against 0.676 at 64); FP64 64 (1.59 ms, 128 the same
# within 1%).
BLOCKED_TILE = {"fp64": 64, "fp32c": 128}
SIC_TH = 0.19669


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def reference_bytes_per_step(net, prec: str) -> float:
    """Bytes one step moves in the reference's data layout, each touched once
    (dense state, one delayed ring slot read per population, every CSR id
    read once). Neuron: 12s state r/w + 2s I_sic + s ring write + s slot read
    + 8 refractory + 12 row_start + 4*ids; astrocyte: 6s state + 4s params +
    s ring write + s slot read + 4 row_start + 4*ids."""
    s = 8 if prec == "fp64" else 4
    n, na = net.n_neurons, net.n_astro
    ids = sum(net.n_edges.values())
    neuron = n * (12 * s + 2 * s + s + s + 8 + 12)
    astro = na * (6 * s + 4 * s + s + s + 4)
    comp = 24 * na if prec == "fp32c" else 0  # Ca/IP3/h compensation terms r/w
    return float(neuron + astro + comp + 4 * ids)


def kernel_bytes_per_launch(net, prec: str, method: str) -> float:
    """Algorithmic HBM bytes of one launch of the method's dominant kernel.

    plain (one step of the gather and cell kernels): the reference layout's
    step (reference_bytes_per_step): the plain method streams every in-edge
    every step.
    blocked (one integrate launch advances every cell one 10-step block): per
    neuron it reads and writes V, w, g_ex, dg_ex, g_in, dg_in, I_sic (14s),
    reads V_reset and b (2s), reads and writes the refractory count (8),
    reads three bucket counts and clears them (24) and writes its spike word
    (4); per astrocyte Ca, IP3, h r/w (6s, plus 6s of compensation terms in
    FP32c), four parameters (4s), bucket count r/w (8), SIC word (4) and the
    ever-active flag (1). Bucket entries, emitted values and STP updates
    scale with activity (a few per cent of cells per block) and are left out.
    """
    if method == "plain":
        return reference_bytes_per_step(net, prec)
    s = 8 if prec == "fp64" else 4
    neuron = 16 * s + 36
    astro = (16 if prec == "fp32c" else 10) * s + 13
    return float(net.n_neurons * neuron + net.n_astro * astro)


def out_degrees(net):
    """Per neuron, the in-CSR entries that deliver its spikes: excitatory and
    neuron-to-astrocyte rows by source id, and inhibitory rows, whose stored
    local ids index the neuron ring as they are (src/astrosim_fused.cu:461)."""
    import numpy as np
    deg = np.zeros(net.n_neurons, dtype=np.int64)
    for which in ("exc", "inh", "e2a"):
        _, src = net.csr(which)
        deg += np.bincount(src, minlength=net.n_neurons)[: net.n_neurons]
        del src
    return deg


def run_leg(A, net, prec: str, method: str, steps: int, warmup: int, block: int, deg) -> dict:
    import numpy as np
    log(f"leg {prec}/{method}: create")
    t0 = time.perf_counter()
    sim = A.Simulation(net, precision=prec, method=method, block=block)
    sim.sync()
    create_s = time.perf_counter() - t0  # device-to-device copy of the network + device setup
    sim.step(warmup)
    sim.sync()
    counts0 = sim.spike_counts() if deg is not None else None
    sim.set_timing(True)
    t0 = time.perf_counter()
    sim.step(steps)
    sim.sync()
    wall = time.perf_counter() - t0
    kms_window, launches_window = sim.kernel_time()
    sim.set_timing(False)
    spikes = sim.spike_count()
    events = None
    if deg is not None:
        d = sim.spike_counts().astype(np.int64) - counts0.astype(np.int64)
        events = int(d @ deg)
    # full launches right after the window: align to a block and run a few
    # untimed launches (the host's readbacks above leave the GPU idle, and its
    # clock ramps down meanwhile), then time whole blocks
    L = 10 if method == "blocked" else 1
    n_full = 20 if method == "blocked" else 100
    sim.step((-sim.steps) % L + (50 if method == "blocked" else 200) * L)
    sim.set_timing(True)
    sim.step(n_full * L)
    kms, launches = sim.kernel_time()
    sim.set_timing(False)
    ca = sim.state("Ca")
    out = {
        "precision": prec, "method": method, "block": block, "steps": steps, "warmup": warmup, "wall_s": wall,
        "steps_per_s": steps / wall, "ms_per_step": 1e3 * wall / steps,
        "spikes_in_window": None if deg is None else int(d.sum()),
        "synaptic_events_per_s": None if events is None else events / wall,
        # the window's kernel time per L steps: a full launch's time, whatever
        # partial launches the window starts or ends with (the basis of the roofline)
        "kernel_ms_per_full_launch": kms_window * L / steps,
        "kernel_ms_window_total": kms_window, "launches_in_window": launches_window,
        "kernel_ms_per_full_launch_after_window": kms / max(1, launches), "full_launches_after_window": launches,
        "steps_per_launch": L, "device_bytes": sim.device_bytes(), "spikes_total": spikes,
        "sic_active_fraction": float(((ca - SIC_TH) * 1000.0 > 1.0).mean()),
        "create_s": create_s,
    }
    sim.close()
    log(f"leg {prec}/{method}: {out['steps_per_s']:.1f} steps/s, {out['kernel_ms_per_full_launch']:.3f} ms "
        f"per full launch" + ("" if events is None else f", {out['synaptic_events_per_s']:.3e} synaptic events/s"))
    return out


def reference_run(A, net, prec: str, pre: int = 1000, sim_steps: int = 10000) -> dict:
    """The reference's run (pre steps, ever-active reset, sim steps) on the
    blocked method: recorded-window rate and SIC fractions
    (src/astrosim_fused.cu:1066-1076, 1117-1129)."""
    import numpy as np
    sim = A.Simulation(net, precision=prec, method="blocked")
    sim.step(pre)
    before = sim.spike_count()
    sim.reset_ever_active()
    sim.step(sim_steps)
    spikes = sim.spike_count() - before
    ca = sim.state("Ca")
    out = {"rate_hz": spikes / (net.n_neurons * sim_steps * 0.1e-3),
           "sic_active": float(np.mean((ca - SIC_TH) * 1000.0 > 1.0)),
           "sic_engaged": float(np.mean(sim.ever_active()))}
    sim.close()
    return out


def accuracy(A, net) -> dict:
    r64 = reference_run(A, net, "fp64")
    r32 = reference_run(A, net, "fp32c")
    out = {"fp64": r64, "fp32c": r32, "rate_rel": r32["rate_hz"] / r64["rate_hz"] - 1.0, "rate_tolerance": 0.01,
           "sic_active_diff": r32["sic_active"] - r64["sic_active"],
           "sic_engaged_diff": r32["sic_engaged"] - r64["sic_engaged"], "sic_tolerance": 0.01,
           "window_steps": 11000,
           "note": "both blocked legs over pre 1,000 + sim 10,000 steps, untimed; FP64 blocked is "
                   "bit-identical to the FP64 host oracle"}
    log(f"fp32c vs fp64 (11,000 steps): rate {100 * out['rate_rel']:+.4f}%, SIC active "
        f"{out['sic_active_diff']:+.6f}, engaged {out['sic_engaged_diff']:+.6f}")
    return out


def cpu_baseline(net, seed: int, n_steps: int) -> dict:
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as O
    csrs = {w: net.csr(w) for w in ("exc", "inh", "sic", "e2a")}
    params = {w: net.params(w) for w in ("Ca_tot", "IP3_0", "tau_IP3", "delta_IP3", "V_reset", "b")}
    onet = O.OracleNet.from_arrays(net.n_total, seed, csrs, params)
    del csrs, params
    threads = min(os.cpu_count() or 1, 16)
    osim = O.OracleSim(onet, threads=threads)
    osim.step(1)  # first touch of the state arrays
    t0 = time.perf_counter()
    osim.step(n_steps)
    wall = time.perf_counter() - t0
    del osim
    # the same stepper on one core, over a few steps
    one = O.OracleSim(onet, threads=1)
    one.step(1)
    n1 = max(1, n_steps // 16)
    t1 = time.perf_counter()
    one.step(n1)
    wall1 = time.perf_counter() - t1
    del one
    model = ""
    try:
        model = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
    except Exception:
        pass
    log(f"cpu baseline: one core {n1 / wall1:.4f} steps/s ({model})")
    return {"value": n_steps / wall, "unit": "steps/s", "cores": threads, "kind": "port",
            "single_core": {"value": n1 / wall1, "steps": n1}, "cpu_model": model,
            "sample": f"FP64 host oracle stepper, {n_steps} steps of the same N_total={net.n_total} "
                      f"network (after 1 warm step), OpenMP over cells"}


def config1_leg(A) -> dict:
    """The reference's default-size run: N_total = 500, seed 1, dt 0.1 ms,
    pre 1000 + sim 5000 ms = 60,000 steps, FP64, launch-bound."""
    net = A.Network(500, 1)
    out = {"workload": "N_total=500 seed=1 dt=0.1ms, 60,000 steps (pre 1000 ms + sim 5000 ms), FP64"}
    for method in ("plain", "blocked"):
        sim = A.Simulation(net, method=method)
        sim.sync()
        t0 = time.perf_counter()
        sim.step(60000)
        sim.sync()
        wall = time.perf_counter() - t0
        out[method] = {"steps_per_s": 60000 / wall, "us_per_step": 1e6 * wall / 60000, "wall_s": wall,
                       "spikes": sim.spike_count()}
        sim.close()
        log(f"config 1 {method}: {60000 / wall:.0f} steps/s ({1e6 * wall / 60000:.2f} us/step)")
    return out


def activity_leg(A, n: int = 1_000_000, seed: int = 7, warm: int = 2000, steps: int = 1000) -> dict:
    """The two FP64 methods on one network at normal activity and at high,
    SIC-driven activity (n2a and a2n weights raised: w_n2a 100, w_a2n 5000).
    The plain method's cost is nearly independent of activity; the blocked
    method's appends grow with the emitters and its buckets overflow into
    whole-row walks. Times `steps` steps after `warm`."""
    from astrosimgpu_cuda_amd import Batch
    net = A.Network(n, seed, setup="gpu")
    out = {"workload": f"N_total={n} seed={seed}, FP64, steps {warm}..{warm + steps}", "cases": []}
    for label, kw in (("normal", {}), ("high", {"w_n2a": 100.0, "w_a2n": 5000.0})):
        case = {"activity": label, **kw}
        sim = A.Simulation(net, method="plain", **kw)
        sim.step(warm)
        sim.sync()
        s0 = sim.spike_count()
        t0 = time.perf_counter()
        sim.step(steps)
        sim.sync()
        wall = time.perf_counter() - t0
        case["plain_steps_per_s"] = steps / wall
        case["rate_hz"] = (sim.spike_count() - s0) / (net.n_neurons * steps * 0.1e-3)
        sim.close()
        b = Batch([net], **kw)
        m = b.add(net, warm + steps)
        b.step(warm)
        b.sync()
        ov0 = b.info()["bucket_overflows"]
        b.set_timing(True)
        t0 = time.perf_counter()
        b.step(steps)
        b.sync()
        wall = time.perf_counter() - t0
        kms, launches = b.kernel_time()
        case["blocked_steps_per_s"] = steps / wall
        case["blocked_integrate_ms_per_block"] = kms / max(1, launches)
        case["blocked_other_ms_per_block"] = 1e3 * wall / max(1, launches) - kms / max(1, launches)
        case["blocked_bucket_overflows_per_block"] = (b.info()["bucket_overflows"] - ov0) / max(1, launches)
        case["blocked_over_plain"] = case["blocked_steps_per_s"] / case["plain_steps_per_s"]
        b.close()
        out["cases"].append(case)
        log(f"activity {label}: {case['rate_hz']:.2f} Hz; plain {case['plain_steps_per_s']:.0f} steps/s, "
            f"blocked {case['blocked_steps_per_s']:.0f} steps/s (integrate "
            f"{case['blocked_integrate_ms_per_block']:.3f} + other {case['blocked_other_ms_per_block']:.3f} ms "
            f"per block, {case['blocked_bucket_overflows_per_block']:.0f} overflows per block)")
    return out


def ensemble_leg(pre_ms: float = 100.0, sim_ms: float = 500.0, n_sample: int = 8) -> dict:
    """Batched ensemble: 256 networks, N_total log-uniform in [500, 200000]
    from seed 3, seeds 1000+i, dt 0.1 ms, pre 100 ms + sim 500 ms, FP32c, in
    one batch (autotuned tile size, hipGraph replays). Compared with running
    networks one at a time: plain FP64 (the reference's method) and blocked
    FP32c, each on an 8-network sample (the size quantiles) extrapolated to
    the 256 networks by the sample's mean time per simulation."""
    from astrosimgpu_cuda_amd import ensemble as E
    steps = int((pre_ms + sim_ms) / 0.1 + 0.5)
    sizes = E.ensemble_sizes(256)
    t0 = time.perf_counter()
    nets = E.build_networks(sizes, setup="gpu")
    build_s = time.perf_counter() - t0
    log(f"ensemble: built 256 networks ({sum(sizes)} cells) in {build_s:.1f} s")
    batched = E.run_batched(nets, steps, precision="fp32c")
    incl = len(nets) / (batched["wall_s"] + batched["setup_s"])
    log(f"ensemble: batched {batched['sims_per_s']:.2f} sims/s ({batched['wall_s']:.2f} s, "
        f"tile {batched['block_threads']}, chunk {batched['chunk_steps']}); with admission {incl:.2f} sims/s")
    order = sorted(range(len(nets)), key=lambda i: sizes[i])
    pick = [order[int((k + 0.5) * len(order) / n_sample)] for k in range(n_sample)]
    sample = [nets[i] for i in pick]
    plain = E.run_one_at_a_time(sample, steps, "fp64", "plain")
    log(f"ensemble: one-at-a-time plain FP64 {plain['sims_per_s']:.2f} sims/s (sample of {n_sample})")
    blk = E.run_one_at_a_time(sample, steps, "fp32c", "blocked")
    log(f"ensemble: one-at-a-time blocked FP32c {blk['sims_per_s']:.2f} sims/s (sample of {n_sample})")
    same = all(batched["spikes"][i] == s for i, s in zip(pick, blk["spikes"]))
    return {
        "networks": 256, "cells": sum(sizes) + sum(max(1, round(n / 5)) for n in sizes),
        "steps_per_sim": steps, "precision": "fp32c",
        "batched_sims_per_s": batched["sims_per_s"], "batched_wall_s": batched["wall_s"],
        "batched_setup_s": batched["setup_s"], "batched_sims_per_s_with_admission": incl,
        "tile_threads": batched["block_threads"], "chunk_steps": batched["chunk_steps"],
        "block_steps": batched["block_steps"], "graph_blocks": batched["graph_blocks"],
        "network_build_s": build_s, "network_setup": "gpu",
        "bucket_overflows": batched["bucket_overflows"], "device_bytes": batched["device_bytes"],
        "one_at_a_time_plain_fp64_sims_per_s": plain["sims_per_s"],
        "one_at_a_time_blocked_fp32c_sims_per_s": blk["sims_per_s"],
        "speedup_vs_plain_fp64": batched["sims_per_s"] / plain["sims_per_s"],
        "speedup_vs_plain_fp64_with_admission": incl / plain["sims_per_s"],
        "speedup_vs_blocked_one_at_a_time": batched["sims_per_s"] / blk["sims_per_s"],
        "sample": f"{n_sample} networks at the size quantiles, N={[sizes[i] for i in pick]}",
        "sampled_members_equal_alone": same,
    }


def serving_leg(duration_s: float, blocks_per_round: int = 16) -> dict:
    """Serving load (40 req/s Poisson arrivals, N_total log-uniform in
    [1e3, 1e6], sim time uniform in [200, 2000] ms, 20% FP64, own seeds)
    replayed in real time for `duration_s` seconds of arrivals against the
    continuous-batching server; latency is submit -> result."""
    from astrosimgpu_cuda_amd import Server, serving
    load = serving.poisson_load(duration_s=duration_s)
    with Server(hbm_fraction=0.5, fp64_share=0.35, build_threads=12, blocks_per_round=blocks_per_round) as srv:
        out = serving.replay(srv, load, timeout_s=900.0)
    st = out["stats"]
    log(f"serving: {st['completed']}/{len(load)} done, p50 {st['p50_latency_s']:.2f} s, "
        f"p99 {st['p99_latency_s']:.2f} s, {st['sims_per_s']:.2f} sims/s")
    res = out["results"]
    build = [r["t_built"] - r["t_submit"] for r in res if r["status"] == "done"]
    queue = [r["t_admit"] - r["t_built"] for r in res if r["status"] == "done"]
    run = [r["t_done"] - r["t_admit"] for r in res if r["status"] == "done"]
    q = lambda v, p: float(sorted(v)[min(len(v) - 1, int(p * len(v)))]) if v else None
    return {
        "arrival_rate_hz": 40.0, "arrival_window_s": duration_s, "requests": len(load),
        "blocks_per_round": blocks_per_round, "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
        "dedicated_queues": bool(st["dedicated_queues"]), "build_streams": st["build_streams"],
        "completed": st["completed"], "rejected": st["rejected"], "all_done": st["all_done"],
        "p50_latency_s": st["p50_latency_s"], "p99_latency_s": st["p99_latency_s"],
        "mean_latency_s": st["mean_latency_s"], "sims_per_s": st["sims_per_s"],
        "hbm_bytes": st["device_bytes"], "peak_active": st["peak_active"],
        "peak_resident_cells": st["peak_resident_cells"], "capacity_cells": st["capacity_cells"],
        "hbm_occupancy_peak": st["peak_resident_cells"] / max(1, st["capacity_cells"]),
        "gpu_busy_s": st["gpu_busy_s"], "cell_steps_per_s": st["cell_steps"] / max(1e-9, st["gpu_busy_s"]),
        "admit_s": st["admit_s"], "retire_s": st["retire_s"],
        "p50_build_s": q(build, 0.5), "p50_queue_s": q(queue, 0.5), "p50_run_s": q(run, 0.5),
        "wall_s": st["wall_s"], "network_setup": "gpu",
    }


def load_profile(kind: str, workload: str) -> dict | None:
    p = ROOT / "profiles" / f"{kind}_{workload}.json"
    if p.exists():
        try:
            return json.loads(p.read_text())
        except Exception:
            return None
    return None


def copy_calibration(nbytes: int) -> int:
    """Three device-to-device copies of `nbytes`: the known-size reference
    that FETCH_SIZE / WRITE_SIZE readings are scaled by (run under
    rocprofv3 --pmc; each copy reads and writes nbytes)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            print("bench.py: hipMalloc failed", file=sys.stderr)
            return 1
    if hip.hipMemset(a, 1, C.c_size_t(nbytes)) != 0:
        return 1
    for _ in range(3):
        if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            return 1
    if hip.hipDeviceSynchronize() != 0:
        return 1
    log(f"copy calibration: {nbytes} bytes copied 3 times")
    hip.hipFree(a)
    hip.hipFree(b)
    return 0


KERNELS = {"plain": "plain_gather_kernel + plain_cell_kernel (one step)",
           "blocked": "integrate_kernel (one 10-step block)"}


def add_roofline(net, leg: dict, workload: str) -> dict:
    """The leg's dominant kernel against the HBM roofline, on the kernel's
    own algorithmic bytes and its full-launch time."""
    prec, method = leg["precision"], leg["method"]
    launch_s = leg["kernel_ms_per_full_launch"] * 1e-3
    own = kernel_bytes_per_launch(net, prec, method)
    achieved = own / launch_s / 1e9
    traffic = load_profile("traffic", f"{method}_{prec}_{workload}")
    valu = load_profile("valu", f"{method}_{prec}_{workload}")
    tb = None if traffic is None else float(traffic["hbm_bytes_per_launch"])
    r = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": achieved / HBM_PEAK_GBS, "traffic": tb,
         "kernel": KERNELS[method], "algorithmic_bytes_per_launch": own,
         "kernel_ms_per_full_launch": leg["kernel_ms_per_full_launch"],
         "measured_gbs": None if tb is None else tb / launch_s / 1e9,
         "measured_frac": None if tb is None else tb / launch_s / 1e9 / HBM_PEAK_GBS,
         "traffic_over_algorithmic": None if tb is None else tb / own,
         "valu_busy_frac": None if valu is None else valu.get("valu_busy_frac"),
         "reference_layout_bytes_per_step": reference_bytes_per_step(net, prec)}
    if method == "blocked":
        r["note"] = ("bound by vector-instruction issue (valu_busy_frac), not by HBM: the kernel reads each "
                     "cell's state once per 10-step block and its inputs from buckets filled by the previous "
                     "block's emitters at the end of their launch")
    else:
        r["note"] = "the plain method streams every in-edge every step: the reference layout's bytes"
    return r


def graph_gaps(trace_csv: str) -> int:
    """Launch gaps of the graph replays in a `rocprofv3 --kernel-trace`
    kernel_trace.csv: the replays are the span from the first to the last
    tick_kernel (a kernel only captured graphs launch, one per block);
    within it, the summed kernel durations against the span are the device
    time a persistent kernel could at most take back. Prints one JSON line."""
    import csv
    rows = []
    with open(trace_csv, newline="") as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    ticks = [i for i, r in enumerate(rows) if "tick_kernel" in r[2]]
    if len(ticks) < 2:
        print(json.dumps({"error": "no graph replays (tick_kernel) in the trace"}))
        return 1
    sel = rows[ticks[0]:ticks[-1] + 1]
    span = sel[-1][1] - sel[0][0]
    busy, end = 0, sel[0][0]
    for a, b, _ in sel:  # union of kernel intervals (kernels of one stream do not overlap)
        a = max(a, end)
        if b > a:
            busy += b - a
        end = max(end, b)
    by = {}
    for a, b, n in sel:
        m = re.search(r"(\w+_kernel|__amd_rocclr_\w+)", n)
        k = m.group(1) if m else n[:40]
        by.setdefault(k, [0, 0])
        by[k][0] += 1
        by[k][1] += b - a
    blocks = len(ticks)
    print(json.dumps({
        "blocks": blocks, "span_ms": span / 1e6, "kernel_busy_ms": busy / 1e6,
        "gap_ms": (span - busy) / 1e6, "gap_fraction": (span - busy) / span,
        "span_us_per_block": span / 1e3 / blocks, "gap_us_per_block": (span - busy) / 1e3 / blocks,
        "kernels": {k: {"calls": c, "avg_us": t / 1e3 / c} for k, (c, t) in by.items()},
    }))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=11000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--plain-steps", type=int, default=0,
                    help="timed steps of the plain FP64 leg (0: the same as --steps)")
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--legs", default="fp64:plain,fp64:blocked,fp32c:blocked")
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--block", type=int, default=128,
                    help="threads per workgroup of the plain method (the CLI's --block)")
    ap.add_argument("--tile", type=int, default=0,
                    help="threads per workgroup (tile) of the blocked method; 0: the fastest measured at N=10M, "
                         "64 for FP64 and 128 for FP32c")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-events", action="store_true", help="skip the synaptic-events count (out-degrees)")
    ap.add_argument("--no-config1", action="store_true")
    ap.add_argument("--no-activity", action="store_true")
    ap.add_argument("--serving-seconds", type=float, default=120.0,
                    help="arrival window of the serving load (0 skips it)")
    ap.add_argument("--serving-blocks", type=int, default=16,
                    help="1 ms blocks the server advances between admission checks")
    ap.add_argument("--graph-gaps", default="", metavar="CSV",
                    help="only analyse the graph replays of a rocprofv3 kernel_trace.csv (no GPU use)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run the known-size copy that HBM counter readings are scaled by")
    args = ap.parse_args()
    if args.graph_gaps:
        return graph_gaps(args.graph_gaps)
    if args.copy_calibration > 0:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); "
              "multi-GPU sharding is a design note, not built", file=sys.stderr)
        return 2

    import astrosimgpu_cuda_amd as A
    name, arch = A.device_info()
    log(f"device {name} ({arch}); building N_total={args.n} seed={args.seed}")
    t0 = time.perf_counter()
    net = A.Network(args.n, args.seed, setup="gpu")
    setup_s = time.perf_counter() - t0
    log(f"network built in {setup_s:.1f} s ({'GPU' if net.built_on_gpu else 'host'} generator, "
        f"connectivity and parameters {net.build_seconds:.2f} s), edges {net.n_edges}")
    deg = None if args.no_events else out_degrees(net)

    legs = [tuple(x.split(":")) for x in args.legs.split(",")]
    results = [run_leg(A, net, p, m, args.plain_steps if (m == "plain" and args.plain_steps > 0) else args.steps,
                       args.warmup, (args.tile or BLOCKED_TILE[p]) if m == "blocked" else args.block, deg)
               for p, m in legs]
    workload = f"n{args.n}_seed{args.seed}"
    for leg in results:
        leg["roofline"] = add_roofline(net, leg, workload)
    head = results[-1]
    by = {(r["precision"], r["method"]): r for r in results}
    ratios = {}
    if ("fp32c", "blocked") in by:
        for key in (("fp64", "blocked"), ("fp64", "plain")):
            if key in by:
                ratios[f"fp32c_blocked_over_{key[0]}_{key[1]}"] = (by[("fp32c", "blocked")]["steps_per_s"]
                                                                  / by[key]["steps_per_s"])
    acc = None if args.no_accuracy else accuracy(A, net)
    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline: FP64 host oracle")
        cpu = cpu_baseline(net, args.seed, args.cpu_steps)
        log(f"cpu baseline: {cpu['value']:.3f} steps/s on {cpu['cores']} threads")
    net.free()
    del deg
    c1 = None if args.no_config1 else config1_leg(A)
    act = None if args.no_activity else activity_leg(A)
    ens = None if args.no_ensemble else ensemble_leg()
    srv = serving_leg(args.serving_seconds, args.serving_blocks) if args.serving_seconds > 0 else None
    line = {
        "metric": "steps/s at N=10M FP32 vs FP64 (1 GPU); ensemble sims/s + p50/p99 latency",
        "value": head["steps_per_s"],
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": head["steps"],
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head["precision"] == "fp64" else "f32+f64",
        "data": "synthetic (network generated from N_total and seed by the reference's scaling law)",
        "config": {"workload": f"single network N_total={args.n} seed={args.seed} dt=0.1ms",
                   "n_total": args.n, "seed": args.seed, "dt_ms": 0.1,
                   "precision": head["precision"], "method": head["method"],
                   "parallelism": "single GPU"},
        "legs": results,
        "speedups": ratios,
        "synaptic_events_per_s": head["synaptic_events_per_s"],
        "network_setup": {"where": "gpu" if net.built_on_gpu else "host", "build_s": net.build_seconds,
                          "wall_s": setup_s},
        "roofline": head["roofline"],
        "cpu_baseline": cpu,
        "fp32c_vs_fp64": acc,
        "config1": c1,
        "activity": act,
        "ensemble": ens,
        "serving": srv,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
