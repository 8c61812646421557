"""BASELINE configurations 3, 4 and 5 on the GPU at their real sizes.

Config 3 (N_total = 10M, seed 42, dt 0.1 ms, pre 100 ms + sim 1000 ms):
  - the FP64 GPU methods against the FP64 host oracle on the same network
    for the first 12 steps: every state array bit-identical. Cells from
    1,000,003 on draw from random streams that alias earlier steps' streams
    (src/astrosim_fused.cu:396, 405, 477, 500), so this covers that region;
  - the plain and blocked FP64 methods bit-identical over 40 steps;
  - FP32c against FP64 (blocked, bit-identical to the oracle) over the
    reference's full 11,000-step run: recorded-window firing rate within 1%,
    fractions of astrocytes above the SIC threshold (final step, and engaged
    during the window) within 0.01.
Config 4: the 256-network ensemble in one batch (FP32c, pre 100 ms + sim
500 ms): every member's spike count and final calcium equal to its solo run.
Config 5: requests drawn from the serving load (40 req/s, N_total
log-uniform in [1e3, 1e6], sim U[200, 2000] ms, 20% FP64), including one of
at least 500,000 cells, replayed against the server with the bench's
settings: every result equal to its solo run.
"""
import numpy as np
import pytest

import astrosimgpu_cuda_amd as A
from astrosimgpu_cuda_amd import ensemble, serving

pytestmark = pytest.mark.gpu

FIELDS = ["Ca", "IP3", "h", "V", "w", "g_ex", "dg_ex", "g_in", "dg_in", "I_sic_hold", "stp_x",
          "stp_u", "stp_t_last"]
SIC_TH = 0.19669


@pytest.fixture(scope="module")
def net10m():
    print("[config3] building N_total=10M seed 42 on the GPU", flush=True)
    net = A.Network(10_000_000, 42, setup="gpu")
    yield net
    net.free()


def differing(a, b):
    return int(np.count_nonzero(a.view(np.uint64) != b.view(np.uint64)))


def test_config3_fp64_matches_oracle_first_steps(net10m):
    import pyoracle as O
    steps = 12
    csrs = {w: net10m.csr(w) for w in ("exc", "inh", "sic", "e2a")}
    params = {w: net10m.params(w) for w in ("Ca_tot", "IP3_0", "tau_IP3", "delta_IP3", "V_reset", "b")}
    onet = O.OracleNet.from_arrays(net10m.n_total, 42, csrs, params)
    del csrs, params
    print(f"[config3] oracle: {steps} steps on the host", flush=True)
    osim = O.OracleSim(onet, threads=0)
    osim.step(steps)
    want = {f: osim.state(f) for f in FIELDS}
    want_spikes = osim.spike_count()
    del osim, onet
    for method in ("plain", "blocked"):
        print(f"[config3] {method} fp64 {steps} steps", flush=True)
        sim = A.Simulation(net10m, method=method)
        sim.step(steps)
        for f in FIELDS:
            assert differing(sim.state(f), want[f]) == 0, (method, f)
        assert sim.spike_count() == want_spikes
        sim.close()


def test_config3_plain_equals_blocked_fp64(net10m):
    sp = A.Simulation(net10m, method="plain")
    sp.step(40)
    got = {f: sp.state(f) for f in FIELDS}
    sps = sp.spike_count()
    sp.close()
    sb = A.Simulation(net10m, method="blocked")
    sb.step(40)
    for f in FIELDS:
        assert differing(got[f], sb.state(f)) == 0, f
    assert sps == sb.spike_count()
    sb.close()


@pytest.mark.timeout(600)
def test_config3_plain_equals_blocked_fp64_through_sic_activity(net10m):
    """12,010 steps: from about step 6,000 astrocytes turn SIC-active (some
    1,800 by the end), neurons then receive SIC from several astrocytes in a
    block, and SIC buckets overflow into whole in-row walks. The blocked
    method (a one-member batch, whose counters report the overflows) stays
    bit-identical to the plain method through all of it."""
    steps = 12010
    print(f"[config3] plain fp64 {steps} steps", flush=True)
    sp = A.Simulation(net10m, method="plain")
    sp.step(steps)
    got = {f: sp.state(f) for f in FIELDS}
    sps = sp.spike_count()
    sp.close()
    print(f"[config3] blocked fp64 {steps} steps", flush=True)
    b = A.Batch([net10m], method="blocked", block=64)
    m = b.add(net10m, steps)
    b.step(steps)
    b.sync()
    for f in FIELDS:
        assert differing(got[f], b.state(m, f)) == 0, f
    assert b.spike_count(m) == sps
    overflows = b.info()["bucket_overflows"]
    active = int(np.count_nonzero((got["Ca"] - SIC_TH) * 1000.0 > 1.0))
    print(f"[config3] spikes {sps}, bucket overflows {overflows}, SIC-active astrocytes {active}", flush=True)
    assert overflows > 0 and active > 100
    b.close()


def run_config3(net, precision):
    pre, sim_steps = 1000, 10000  # pre 100 ms, sim 1000 ms at dt 0.1 ms
    sim = A.Simulation(net, precision=precision, method="blocked")
    sim.step(pre)
    before = sim.spike_count()
    sim.reset_ever_active()
    sim.step(sim_steps)
    spikes = sim.spike_count() - before
    ca = sim.state("Ca")
    out = {"rate_hz": spikes / (net.n_neurons * 1.0), "active": float(np.mean((ca - SIC_TH) * 1000.0 > 1.0)),
           "engaged": float(np.mean(sim.ever_active()))}
    sim.close()
    return out


def test_config3_fp32c_full_run_statistics(net10m):
    r64 = run_config3(net10m, "fp64")
    r32 = run_config3(net10m, "fp32c")
    print(f"[config3] fp64 {r64} fp32c {r32}", flush=True)
    assert r64["rate_hz"] > 0
    assert abs(r32["rate_hz"] / r64["rate_hz"] - 1.0) < 0.01
    assert abs(r32["active"] - r64["active"]) < 0.01
    assert abs(r32["engaged"] - r64["engaged"]) < 0.01


def test_config4_ensemble_members_equal_solo_runs():
    steps = int((100.0 + 500.0) / 0.1 + 0.5)
    sizes = ensemble.ensemble_sizes(256)
    print(f"[config4] building 256 networks ({sum(sizes)} neurons)", flush=True)
    nets = ensemble.build_networks(sizes, setup="gpu")
    batch = A.Batch(nets, precision="fp32c")
    members = [batch.add(n, steps) for n in nets]
    L = batch.info()["block_steps"]
    chosen = batch.autotune((64, 128, 256), 2)
    done = 3 * 2 * L
    while done + 50 * L <= steps:
        batch.run_graph(50)
        done += 50 * L
    batch.step(steps - done)
    batch.sync()
    print(f"[config4] batch done (tile {chosen}); solo runs", flush=True)
    got = [(batch.spike_count(m), batch.state(m, "Ca")) for m in members]
    batch.close()
    for i, net in enumerate(nets):
        sim = A.Simulation(net, precision="fp32c", method="blocked")
        sim.step(steps)
        assert sim.spike_count() == got[i][0], (i, sizes[i])
        assert differing(sim.state("Ca"), got[i][1]) == 0, (i, sizes[i])
        sim.close()


def solo(n_total, seed, steps, precision):
    net = A.Network(n_total, seed, setup="gpu")
    sim = A.Simulation(net, method="blocked", precision=precision)
    sim.step(steps)
    ca = sim.state("Ca")
    out = (sim.spike_count(), ca.mean(), int(np.sum((ca - SIC_TH) * 1000.0 > 1.0)))
    sim.close()
    return out


def test_config5_served_results_equal_solo_runs():
    full = serving.poisson_load(duration_s=120.0)
    load = full[:36]
    big = next(r for r in full if r["n_total"] >= 500_000)
    if big not in load:
        load = load + [dict(big, t=load[-1]["t"] + 0.025)]
    n64 = sum(r["precision"] == "fp64" for r in load)
    print(f"[config5] {len(load)} requests over {load[-1]['t']:.2f} s, {n64} FP64, largest "
          f"{max(r['n_total'] for r in load)} cells", flush=True)
    assert max(r["n_total"] for r in load) >= 500_000 and n64 >= 3
    with A.Server(hbm_fraction=0.5, fp64_share=0.35, build_threads=12, blocks_per_round=16) as srv:
        out = serving.replay(srv, load, timeout_s=240.0)
    st = out["stats"]
    print(f"[config5] p50 {st['p50_latency_s']:.2f} s p99 {st['p99_latency_s']:.2f} s, peak resident "
          f"{st['peak_active']}", flush=True)
    assert st["all_done"] and st["completed"] == len(load) and st["rejected"] == 0
    for req, res in zip(load, out["results"]):
        spikes, mean_ca, active = solo(req["n_total"], req["seed"], req["n_steps"], req["precision"])
        assert res["spikes"] == spikes, (req, res)
        assert res["mean_ca"] == pytest.approx(mean_ca, rel=1e-12), (req, res)
        assert res["active_astro"] == active, (req, res)
    assert st["peak_active"] >= 2
