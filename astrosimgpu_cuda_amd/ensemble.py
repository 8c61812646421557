"""Batched ensembles: many independent networks of uneven size advanced
together in one batch (one integrate launch per 1 ms block,
replayed from a captured hipGraph), with the workgroup/tile size chosen by the
batch autotuner. The reference runs one network per process; running the same
networks one at a time is the comparison.
"""
from __future__ import annotations

import math
import time
from concurrent.futures import ThreadPoolExecutor

from . import Batch, Network, Simulation

MASK64 = (1 << 64) - 1


def splitmix_uniform(seed: int, stream: int, counter: int) -> float:
    """The counter RNG's uniform draw (src/astrosim_fused.cu:127-141)."""
    z = (seed + stream * 0x9E3779B97F4A7C15 + counter * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    z ^= z >> 31
    return (z >> 11) * (1.0 / 9007199254740992.0)


def ensemble_sizes(n: int = 256, lo: int = 500, hi: int = 200_000, seed: int = 3) -> list[int]:
    """N_total of network i: log-uniform in [lo, hi] from draw i of stream 0 of `seed`."""
    a, b = math.log(lo), math.log(hi)
    return [int(round(math.exp(a + splitmix_uniform(seed, 0, i) * (b - a)))) for i in range(n)]


def build_networks(sizes: list[int], seed0: int = 1000, workers: int = 8, setup: str = "host") -> list[Network]:
    """Builds network i with seed seed0 + i (the C calls run in parallel);
    ``setup`` as for Network ("gpu": the parallel generator on the GPU)."""
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(lambda a: Network(a[1], seed0 + a[0], setup=setup), enumerate(sizes)))


def run_batched(nets: list[Network], steps: int, precision: str = "fp32c", graph_blocks: int = 50,
                autotune: tuple[int, ...] | None = (64, 128, 256, 512), tune_blocks: int = 2) -> dict:
    """Runs every network `steps` steps in one batch; returns timings and the
    per-network spike counts."""
    t_create = time.perf_counter()
    batch = Batch(nets, precision=precision)
    members = batch.add_many(nets, steps)
    batch.sync()
    t0 = time.perf_counter()
    L = batch.info()["block_steps"]
    done = 0
    chosen = batch.info()["block_threads"]
    if autotune:
        # runs tune_blocks real blocks per (workgroup size, LDS chunk) pair
        chosen = batch.autotune(autotune, tune_blocks)
        done = batch.info()["global_step"]
    while done + graph_blocks * L <= steps:
        batch.run_graph(graph_blocks)
        done += graph_blocks * L
    if done < steps:
        batch.step(steps - done)
    batch.sync()
    wall = time.perf_counter() - t0
    spikes = [batch.spike_count(m) for m in members]
    info = batch.info()
    batch.close()
    return {"wall_s": wall, "setup_s": t0 - t_create, "sims": len(nets), "sims_per_s": len(nets) / wall,
            "steps": steps, "block_threads": chosen, "chunk_steps": info["chunk_steps"], "block_steps": L,
            "graph_blocks": graph_blocks, "device_bytes": info["device_bytes"],
            "bucket_overflows": info["bucket_overflows"], "spikes": spikes}


def run_one_at_a_time(nets: list[Network], steps: int, precision: str, method: str) -> dict:
    """Runs the networks one after another, each alone on the GPU."""
    walls, spikes = [], []
    for n in nets:
        sim = Simulation(n, precision=precision, method=method)
        sim.sync()
        t0 = time.perf_counter()
        sim.step(steps)
        sim.sync()
        walls.append(time.perf_counter() - t0)
        spikes.append(sim.spike_count())
        sim.close()
    total = sum(walls)
    return {"wall_s": total, "sims": len(nets), "sims_per_s": len(nets) / total, "per_sim_s": walls,
            "spikes": spikes}
