"""The continuous-batching server: requests arriving while others run are
admitted as capacity frees up, and each result is bit-identical to running
the same network alone (FP32c and FP64)."""
import time

import numpy as np
import pytest

import astrosimgpu_cuda_amd as A
from astrosimgpu_cuda_amd import serving

pytestmark = pytest.mark.gpu


def solo(n_total, seed, steps, precision):
    net = A.Network(n_total, seed)
    sim = A.Simulation(net, method="blocked", precision=precision)
    sim.step(steps)
    ca = sim.state("Ca")
    out = (sim.spike_count(), ca.mean(), int(np.sum((ca - 0.19669) * 1000.0 > 1.0)))
    sim.close()
    return out


# (blocks per round, largest N_total, arena share of free HBM, load seed):
# short rounds with host-built networks; the default 16-block rounds with
# four networks of 125k-148k cells generated on the GPU (one FP64)
CASES = [(2, 30000, 0.002, 9), (16, 150000, 0.02, 13)]


@pytest.mark.parametrize("blocks,n_hi,hbm,seed", CASES)
def test_server_results_equal_solo_runs(blocks, n_hi, hbm, seed):
    # a near-simultaneous burst, so requests arrive while others run
    load = serving.poisson_load(rate_hz=400.0, duration_s=0.05, n_lo=1000, n_hi=n_hi, sim_lo_ms=100.0,
                                sim_hi_ms=300.0, fp64_fraction=0.3, seed=seed)
    print(f"[server] {len(load)} requests, {blocks} blocks per round", flush=True)
    assert len(load) >= 8
    # a small arena so that later requests must wait for earlier ones to leave
    with A.Server(hbm_fraction=hbm, fp64_share=0.3, build_threads=4, blocks_per_round=blocks) as srv:
        out = serving.replay(srv, load, timeout_s=240.0)
    st = out["stats"]
    print(f"[server] stats {st}", flush=True)
    assert st["all_done"] and st["completed"] == len(load) and st["rejected"] == 0
    # every stream on a hardware queue of its own: two block streams, one
    # upload stream and one build stream per builder (4)
    assert st["dedicated_queues"] == 1 and st["build_streams"] == 4
    n_gpu = sum(r["built_on_gpu"] for r in out["results"])
    print(f"[server] {n_gpu} of {len(load)} networks generated on the GPU", flush=True)
    if n_hi >= 100000:
        assert n_gpu >= 1
    for req, res in zip(load, out["results"]):
        assert res["status"] == "done"
        assert res["t_submit"] <= res["t_build_start"] <= res["t_built"] <= res["t_admit"] <= res["t_done"]
        spikes, mean_ca, active = solo(req["n_total"], req["seed"], req["n_steps"], req["precision"])
        assert res["spikes"] == spikes, (req, res)
        # the calcium values are bit-identical; only the summation of the mean differs
        assert res["mean_ca"] == pytest.approx(mean_ca, rel=1e-12), (req, res)
        assert res["active_astro"] == active, (req, res)
    assert st["peak_active"] >= 2  # requests overlapped on the device


def test_driver_failure_ends_the_wait_with_the_error():
    # the fail_after_rounds hook makes the driver raise after 3 rounds (no
    # kernel is involved: the error is thrown on the host between rounds)
    with A.Server(hbm_fraction=0.01, fp64_share=0.3, build_threads=2, blocks_per_round=2,
                  fail_after_rounds=3) as srv:
        ids = [srv.submit(2000 + 100 * i, 50 + i, 3000, "fp32c") for i in range(6)]
        t0 = time.perf_counter()
        with pytest.raises(A.AstrosimError, match="injected failure"):
            srv.wait_all(120.0)
        assert time.perf_counter() - t0 < 60.0
        statuses = {srv.result(i)["status"] for i in ids}
        assert statuses <= {"done", "rejected"} and "rejected" in statuses
        with pytest.raises(A.AstrosimError, match="injected failure"):
            srv.submit(2000, 1, 100, "fp32c")
        with pytest.raises(A.AstrosimError):
            srv.stats()
