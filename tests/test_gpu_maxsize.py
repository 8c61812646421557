"""The largest network the library accepts: N_total = 40M, seed 7.

32M excitatory + 8M inhibitory neurons and 8M astrocytes; the excitatory
in-CSR holds more than 2^31 edges, so every edge offset past the signed
32-bit range is exercised, and cell ids run to 40M (streams aliasing from
cell 1,000,003 on, src/astrosim_fused.cu:396, 405, 477, 500).
  - FP64 plain and blocked against the FP64 host oracle on the same network
    over 40 steps (four 10-step blocks; the first spikes come from step 21
    on, so spike delivery through both methods is covered): every state
    array bit-identical, the same spike count;
  - FP32c blocked over the same 40 steps: finite state, spike count within
    2% of FP64.
Sizes beyond 40M are rejected (the edge ids are 32-bit). The parallel
generator declines CSRs of 2^31 edges or more (its sorts count in int), so
this network comes from the host walk and is staged to the device.
"""
import numpy as np
import pytest

import astrosimgpu_cuda_amd as A

pytestmark = pytest.mark.gpu

N_MAX = 40_000_000
SEED = 7
FIELDS = ["Ca", "IP3", "h", "V", "w", "g_ex", "dg_ex", "g_in", "dg_in", "I_sic_hold", "stp_x",
          "stp_u", "stp_t_last"]


def differing(a, b):
    return int(np.count_nonzero(a.view(np.uint64) != b.view(np.uint64)))


@pytest.fixture(scope="module")
def net40m():
    print(f"[maxsize] building N_total={N_MAX} seed {SEED} on the GPU", flush=True)
    net = A.Network(N_MAX, SEED, setup="gpu")
    where = "GPU" if net.built_on_gpu else "host"
    print(f"[maxsize] built on the {where} in {net.build_seconds:.2f} s, edges {net.n_edges}", flush=True)
    yield net
    net.free()


def fits(net, **kw):
    free, _ = A.device_memory()
    return A.Simulation.footprint(net, **kw) < free


def test_maxsize_shape(net40m):
    assert (net40m.n_exc, net40m.n_inh, net40m.n_astro) == (32_000_000, 8_000_000, 8_000_000)
    assert net40m.n_edges["exc"] > 2**31
    with pytest.raises(A.AstrosimError):
        A.Network(N_MAX + 1, SEED, setup="gpu")


@pytest.mark.timeout(900)
def test_maxsize_matches_oracle(net40m):
    import pyoracle as O
    steps = 40
    kws = ({"method": "plain"}, {"method": "blocked"}, {"method": "blocked", "precision": "fp32c"})
    for kw in kws:
        if not fits(net40m, **kw):
            pytest.skip(f"{kw} at 40M does not fit the free device memory")
    csrs = {w: net40m.csr(w) for w in ("exc", "inh", "sic", "e2a")}
    params = {w: net40m.params(w) for w in ("Ca_tot", "IP3_0", "tau_IP3", "delta_IP3", "V_reset", "b")}
    onet = O.OracleNet.from_arrays(net40m.n_total, SEED, csrs, params)
    del csrs, params
    print(f"[maxsize] oracle: {steps} steps on the host", flush=True)
    osim = O.OracleSim(onet, threads=0)
    osim.step(steps)
    want = {f: osim.state(f) for f in FIELDS}
    want_spikes = osim.spike_count()
    del osim, onet
    print(f"[maxsize] oracle spikes {want_spikes}", flush=True)
    assert want_spikes > 0
    for kw in kws[:2]:
        print(f"[maxsize] {kw['method']} fp64 {steps} steps", flush=True)
        sim = A.Simulation(net40m, **kw)
        sim.step(steps)
        for f in FIELDS:
            assert differing(sim.state(f), want[f]) == 0, (kw, f)
        assert sim.spike_count() == want_spikes
        sim.close()
    print(f"[maxsize] blocked fp32c {steps} steps", flush=True)
    s32 = A.Simulation(net40m, **kws[2])
    s32.step(steps)
    spikes32 = s32.spike_count()
    for f in ("Ca", "IP3", "h", "V", "w", "g_ex", "g_in"):
        assert np.all(np.isfinite(s32.state(f))), f
    s32.close()
    print(f"[maxsize] spikes fp64 {want_spikes} fp32c {spikes32}", flush=True)
    assert abs(spikes32 - want_spikes) <= 0.02 * want_spikes
