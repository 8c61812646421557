"""Serving load: simulation requests arriving as a Poisson process, replayed
against the continuous-batching server in real time."""
from __future__ import annotations

import math
import time

from . import Server
from .ensemble import splitmix_uniform


def poisson_load(rate_hz: float = 40.0, duration_s: float = 120.0, n_lo: int = 1000, n_hi: int = 1_000_000,
                 sim_lo_ms: float = 200.0, sim_hi_ms: float = 2000.0, fp64_fraction: float = 0.2,
                 dt: float = 0.1, seed: int = 5) -> list[dict]:
    """Requests of the serving load. Arrival gaps are exponential with mean
    1/rate (stream 0 of `seed`), N_total log-uniform in [n_lo, n_hi] (stream
    1), simulated time uniform in [sim_lo_ms, sim_hi_ms] (stream 2), FP64 for
    a `fp64_fraction` share (stream 3); request i gets seed 100000 + i."""
    out, t, i = [], 0.0, 0
    a, b = math.log(n_lo), math.log(n_hi)
    while True:
        t += -math.log(1.0 - splitmix_uniform(seed, 0, i)) / rate_hz
        if t >= duration_s:
            return out
        n = int(round(math.exp(a + splitmix_uniform(seed, 1, i) * (b - a))))
        sim_ms = sim_lo_ms + splitmix_uniform(seed, 2, i) * (sim_hi_ms - sim_lo_ms)
        prec = "fp64" if splitmix_uniform(seed, 3, i) < fp64_fraction else "fp32c"
        out.append({"t": t, "n_total": n, "seed": 100000 + i, "n_steps": int(sim_ms / dt + 0.5),
                    "precision": prec})
        i += 1


def replay(server: Server, load: list[dict], timeout_s: float = 600.0) -> dict:
    """Submits each request at its arrival time, waits for all, returns the
    server statistics and per-request results."""
    t0 = time.perf_counter()
    ids = []
    for r in load:
        delay = r["t"] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append(server.submit(r["n_total"], r["seed"], r["n_steps"], r["precision"]))
    done = server.wait_all(timeout_s)
    wall = time.perf_counter() - t0
    stats = server.stats()
    stats["all_done"] = done
    stats["wall_s"] = wall
    return {"stats": stats, "results": [server.result(i) for i in ids]}
