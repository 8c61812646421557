"""Spike recording through the C ABI (as_sim_read_spikes): the recorded
events equal the FP64 host oracle's raster over the same steps, per-neuron
totals equal the device counters, and the calcium record sampled every step
gives exactly the reference's engaged count (ever_active,
src/astrosim_fused.cu:423, 1121-1125)."""
import numpy as np
import pytest

import astrosimgpu_cuda_amd as A
import pyoracle as O
from astrosimgpu_cuda_amd import recording as R

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method", ["plain", "blocked"])
def test_recording_matches_oracle_raster_and_engaged_count(method, tmp_path):
    n, seed, steps = 2000, 7, 1600
    kw = {"w_n2a": 100.0, "w_a2n": 5000.0}  # astrocytes reach the SIC threshold in this window
    print(f"[record] {method} n={n} steps={steps}", flush=True)
    net, onet = A.Network(n, seed), O.OracleNet(n, seed)
    sim = A.Simulation(net, method=method, **kw)
    meta = R.record(sim, steps, tmp_path / "rec", ca_every=1)
    rec = R.load(tmp_path / "rec")
    osim = O.OracleSim(onet, threads=0, **kw)
    raster = osim.step(steps, raster=True)
    want_steps, want_neurons = np.nonzero(raster)
    assert meta["n_spikes"] == len(want_steps) > 0
    assert np.array_equal(rec["spike_steps"], want_steps)
    assert np.array_equal(rec["spike_neurons"], want_neurons)
    assert np.array_equal(np.bincount(rec["spike_neurons"], minlength=n), sim.spike_counts())
    t = R.transients(rec["ca"])
    assert t["engaged"] == int(sim.ever_active().sum()) > 0
    assert np.array_equal(rec["ca"][-1], osim.state("Ca"))


def test_read_spikes_cap_and_range():
    net = A.Network(20000, 5)
    sim = A.Simulation(net, method="plain", w_n2a=100.0, w_a2n=5000.0)
    s_all = []
    while len(s_all) <= 10 and sim.steps < 5000:  # a window with spikes to truncate
        sim.step(10)
        s_all, n_all = sim.spikes(sim.steps - 9, 9)
    first = sim.steps - 9
    s3, n3 = sim.spikes(first, 9, cap=3)
    assert len(s_all) > 10
    assert np.array_equal(s3, s_all[:3]) and np.array_equal(n3, n_all[:3])
    with pytest.raises(A.AstrosimError):
        sim.spikes(10, 5)  # no longer held
