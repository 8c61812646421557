"""GPU FP64 path against the FP64 host oracle: identical spike raster step
by step and bit-identical state (reference: src/astrosim_fused.cu:374-587,
1042-1064). Calls go through the C ABI (libastrosim.so)."""
import numpy as np
import pytest

import astrosimgpu_cuda_amd as A
import pyoracle as O

pytestmark = pytest.mark.gpu

FIELDS = ["Ca", "IP3", "h", "V", "w", "g_ex", "dg_ex", "g_in", "dg_in", "I_sic_hold", "stp_x",
          "stp_u", "stp_t_last"]


def assert_state_bitwise(sim, osim, where=""):
    for f in FIELDS:
        g, o = sim.state(f), osim.state(f)
        bad = np.nonzero(g.view(np.uint64) != o.view(np.uint64))[0]
        assert bad.size == 0, f"{where} {f}: {bad.size} cells differ, first {bad[:5]} gpu={g[bad[:3]]} oracle={o[bad[:3]]}"
    assert np.array_equal(sim.refractory(), osim.refractory())
    assert np.array_equal(sim.ever_active(), osim.ever_active())
    assert sim.spike_count() == osim.spike_count()


def run_pair(n, seed, steps, chunk=10, method="plain", oracle_kw=None, **kw):
    print(f"[parity] n={n} seed={seed} steps={steps} method={method} {kw}", flush=True)
    net, onet = A.Network(n, seed), O.OracleNet(n, seed)
    sim = A.Simulation(net, method=method, **kw)
    okw = {k: kw[k] for k in ("dt", "substeps", "w_n2a", "w_a2n", "stp") if k in kw}
    osim = O.OracleSim(onet, threads=0, **okw)
    done = 0
    per_neuron = np.zeros(net.n_neurons, dtype=np.int64)
    while done < steps:
        k = min(chunk, steps - done)
        sim.step(k)
        want = osim.step(k, raster=True)
        got = sim.raster(done, k)
        if not np.array_equal(got, want):
            s, j = np.argwhere(got != want)[0]
            pytest.fail(f"raster differs at step {done + s} neuron {j}: gpu={got[s, j]} oracle={want[s, j]}")
        per_neuron += want.sum(axis=0, dtype=np.int64)
        done += k
    assert_state_bitwise(sim, osim, f"after {steps} steps")
    # per-neuron spike counts: the oracle's raster summed over every step
    assert np.array_equal(sim.spike_counts().astype(np.int64), per_neuron)
    return sim, osim


METHODS = ["plain", "blocked"]


@pytest.mark.parametrize("method", METHODS)
def test_config1_first_100ms_bit_exact(method):
    # N_total=500, seed 1, dt 0.1: the first 100 ms (1000 steps)
    run_pair(500, 1, 1000, method=method)


@pytest.mark.parametrize("method", METHODS)
def test_overrides_no_stp_bit_exact(method):
    run_pair(3000, 7, 400, w_n2a=0.5, w_a2n=2.0, stp=False, method=method)


@pytest.mark.parametrize("method", METHODS)
def test_substeps_and_dt_bit_exact(method):
    # dt 0.05 ms: delays of 20 steps
    run_pair(1000, 3, 400, dt=0.05, substeps=2, chunk=20, method=method)


def test_blocked_coarse_dt_bit_exact():
    # dt 0.3 ms: delays of 3 steps, blocks of 3
    run_pair(2000, 4, 300, dt=0.3, chunk=3, method="blocked")


@pytest.mark.parametrize("method", METHODS)
def test_strong_poisson_drive_bit_exact(method):
    # dt 1 ms: Poisson means 2.7 (excitatory), 2.5 (inhibitory) and 3e-3
    # (astrocytes) per step, so most steps draw several uniforms and the
    # blocked method's merged Knuth continuation runs long per-lane chains;
    # the delay is one step (one-step blocks; the raster ring holds 2 steps)
    run_pair(2000, 11, 200, dt=1.0, chunk=2, method=method)


def test_blocked_unaligned_chunks_bit_exact():
    # launches that start and end inside blocks (partial blocks)
    run_pair(4000, 9, 301, chunk=7, method="blocked")


@pytest.mark.parametrize("method,block", [("plain", 64), ("plain", 1024), ("blocked", 64),
                                          ("blocked", 256), ("blocked", 1024)])
def test_block_size_does_not_change_results(method, block):
    run_pair(700, 5, 200, block=block, method=method)


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("n", [1, 7, 20])
def test_tiny_networks(method, n):
    run_pair(n, 2, 300, method=method)


def test_plain_split_launches_bit_exact():
    # from 2^16 neurons the plain method runs a gather kernel and a cell
    # kernel per step (summary bits in LDS confirmed against the flag words)
    run_pair(70000, 13, 300, method="plain", chunk=10)


@pytest.mark.parametrize("method", METHODS)
def test_poisson_normal_approximation_bit_exact(method):
    # dt 12 ms: the excitatory Poisson mean is 2700 Hz * 12 ms = 32.4 > 30,
    # so rng_poisson takes its normal-approximation branch (src:160-165);
    # the inhibitory mean is exactly 30 (Knuth, long chains); one-step delays
    run_pair(3000, 17, 60, dt=12.0, chunk=1, method=method)


def test_blocked_larger_network_bit_exact():
    # 50k cells, 200 ms: exercises bucket overflow fallbacks at the onset
    run_pair(50000, 11, 2000, method="blocked", chunk=10)


@pytest.mark.parametrize("method", METHODS)
def test_high_activity_bit_exact(method):
    # Strong n2a/a2n weights: normal activity for ~150 ms, then SIC-driven
    # firing at up to one spike per step. The blocked method crosses every
    # input path: attached single-spike entries, sources that spiked several
    # times in a block (values read from the ring), more entries than fit in
    # registers, and overflowed buckets (the whole in-row).
    run_pair(2000, 7, 2000, w_n2a=100.0, w_a2n=5000.0, method=method)


@pytest.mark.parametrize("method", METHODS)
def test_config1_full_run_summary_matches(method):
    # pre 1000 ms + sim 5000 ms as the reference's default run
    net, onet = A.Network(500, 1), O.OracleNet(500, 1)
    sim, osim = A.Simulation(net, method=method), O.OracleSim(onet, threads=0)
    sim.step(10000)
    osim.step(10000)
    assert sim.spike_count() == osim.spike_count()
    sim.reset_ever_active()
    osim.reset_ever_active()
    sim.step(50000)
    osim.step(50000)
    assert_state_bitwise(sim, osim, "config 1 end")


@pytest.mark.parametrize("method", METHODS)
def test_cli_summary_matches_oracle_cli(method):
    import subprocess
    args = ["--n", "500", "--seed", "1", "--pre-ms", "200", "--sim-ms", "800"]
    g = A.run_cli(args + ["--method", method], timeout=300)
    assert g.returncode == 0, g.stderr
    o = subprocess.run([str(O.BIN), *args], capture_output=True, text=True, timeout=300)
    pick = lambda out, key: [l for l in out.splitlines() if l.startswith(key)]
    for key in ("spikes (recorded win)", "mean firing rate", "mean calcium", "astrocytes active",
                "astrocytes engaged", "connectivity: exc="):
        gl, ol = pick(g.stdout, key), pick(o.stdout, key)
        if key.startswith("connectivity"):
            gl, ol = [l.split(", built")[0] for l in gl], [l.split(", built")[0] for l in ol]
        assert gl == ol, (key, gl, ol)
    assert "windowed rate" in g.stdout
    assert "GPU: " in g.stdout
