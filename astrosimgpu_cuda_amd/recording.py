"""Spike and calcium recording to disk, and calcium-transient analysis.

The reference prints only summary figures; recording spikes and calcium and
analysing transients are left to the upstream tools it names
(README.md:120-124). Here a :class:`Simulation` is advanced in chunks of at
most one minimum delay (the steps every method still holds), and after each
chunk:
  - every spike of the chunk is read as (step, neuron) events, compacted on
    the device (``Simulation.spikes``, the C ABI's ``as_sim_read_spikes``);
  - every ``ca_every`` steps the astrocytes' calcium is sampled.
Files in the output directory (raw little-endian arrays, appended chunk by
chunk, and a JSON header):
  spike_steps.i64, spike_neurons.u32  the events, in step then neuron order
  ca_steps.i64, ca.f64                sample steps, and [samples][n_astro] calcium
  meta.json                           sizes, dt, window, sampling period
:func:`load` maps them back; :func:`transients` counts calcium transients
(upward crossings of the reference's SIC activity test, (Ca - SIC_th)*1000 >
1, src/astrosim_fused.cu:416-417, 1117) and the astrocytes engaged at least
once -- with ``ca_every=1`` exactly the reference's "engaged" count
(``ever_active``, src/astrosim_fused.cu:423, 1121-1125).
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np

SIC_TH = 0.19669  # src/astrosim_fused.cu:746


def min_delay_steps(dt: float) -> int:
    """Steps of the 1 ms synaptic delay (src/astrosim_fused.cu:783-786): the
    steps every method holds after a chunk."""
    return max(1, int(1.0 / dt + 0.5))


def record(sim, n_steps: int, out_dir, ca_every: int = 10, dt: float | None = None,
           no_alloc: bool = False) -> dict:
    """Advances ``sim`` by ``n_steps`` steps, writing its spikes and sampled
    calcium to ``out_dir``. Returns the header written to meta.json."""
    dt = sim.cfg.dt if dt is None else dt
    out = Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    chunk = min_delay_steps(dt)
    first = sim.steps
    files = {name: open(out / name, "wb") for name in ("spike_steps.i64", "spike_neurons.u32", "ca_steps.i64", "ca.f64")}
    n_spikes = n_samples = 0
    try:
        done = 0
        while done < n_steps:
            # chunks end on calcium sample steps so a sample follows its chunk
            k = min(chunk, n_steps - done)
            if ca_every > 0:
                to_sample = ca_every - ((sim.steps - first) % ca_every)
                k = min(k, to_sample)
            s0 = sim.steps
            sim.step(k)
            steps, neurons = sim.spikes(s0, k, no_alloc=no_alloc)
            files["spike_steps.i64"].write(steps.astype("<i8").tobytes())
            files["spike_neurons.u32"].write(neurons.astype("<u4").tobytes())
            n_spikes += len(steps)
            done += k
            if ca_every > 0 and (sim.steps - first) % ca_every == 0:
                files["ca_steps.i64"].write(np.array([sim.steps], dtype="<i8").tobytes())
                files["ca.f64"].write(sim.state("Ca").astype("<f8").tobytes())
                n_samples += 1
    finally:
        for f in files.values():
            f.close()
    meta = {"n_neurons": sim.n_neurons, "n_astro": sim.n_astro, "dt_ms": dt, "first_step": first,
            "n_steps": n_steps, "ca_every": ca_every, "n_spikes": n_spikes, "n_ca_samples": n_samples}
    (out / "meta.json").write_text(json.dumps(meta, indent=1) + "\n")
    return meta


def load(out_dir) -> dict:
    """The recording as arrays (memory-mapped where non-empty)."""
    d = Path(out_dir)
    meta = json.loads((d / "meta.json").read_text())

    def arr(name, dtype, shape=None):
        p = d / name
        if p.stat().st_size == 0:
            return np.zeros((0,) if shape is None else (0, shape[1]), dtype=dtype)
        a = np.memmap(p, dtype=dtype, mode="r")
        return a if shape is None else a.reshape(shape)

    return {"meta": meta,
            "spike_steps": arr("spike_steps.i64", "<i8"),
            "spike_neurons": arr("spike_neurons.u32", "<u4"),
            "ca_steps": arr("ca_steps.i64", "<i8"),
            "ca": arr("ca.f64", "<f8", (meta["n_ca_samples"], meta["n_astro"]))}


def transients(ca: np.ndarray, sic_th: float = SIC_TH) -> dict:
    """Calcium transients in a [samples][astrocytes] calcium record: an
    astrocyte is active in a sample when (Ca - SIC_th)*1000 > 1 (the
    reference's test); a transient starts at each inactive -> active change
    (and at the first sample if already active). Returns per-astrocyte
    transient counts, the total, the engaged count (active in any sample) and
    the mean transient length in samples."""
    ca = np.asarray(ca)
    if ca.ndim != 2 or ca.shape[0] == 0:
        n = ca.shape[1] if ca.ndim == 2 else 0
        return {"per_astro": np.zeros(n, dtype=np.int64), "total": 0, "engaged": 0, "mean_length": math.nan}
    active = (ca - sic_th) * 1000.0 > 1.0
    starts = active.copy()
    starts[1:] &= ~active[:-1]
    per = starts.sum(axis=0).astype(np.int64)
    total = int(per.sum())
    return {"per_astro": per, "total": total, "engaged": int(active.any(axis=0).sum()),
            "mean_length": float(active.sum()) / total if total else math.nan}
