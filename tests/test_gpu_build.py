"""The connectivity generated on the GPU (as_network_build_on, csrc/
as_gpubuild.hip) against the sequential host walk of the setup stream
(src/astrosim_fused.cu:631-735): the four CSRs and the per-cell parameters
must be identical bit for bit, from the smallest network the generator
accepts to a million cells."""
import numpy as np
import pytest

import astrosimgpu_cuda_amd as A
import selftest_math as S

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_total,seed", [(2000, 1), (20000, 3), (145000, 42), (1000000, 7), (10000000, 42)])
def test_gpu_build_equals_host_build(n_total, seed):
    print(f"gpu build compare N={n_total} seed={seed}", flush=True)
    rc, st = S.build_compare(n_total, seed, gpu=True)
    print(f"  skips decided by the C library: {st['ambiguous']:.0f}, autapse repairs: "
          f"{st['repairs0']:.0f} + {st['repairs2']:.0f}", flush=True)
    assert rc == 0, (rc, st)


# random sizes and seeds (one of them has a skip the C library decides)
@pytest.mark.parametrize("n_total,seed", [(465821, 601051018), (1843701, 838826498), (136123, 869349200),
                                          (1692482, 67822819), (110773, 498145836), (547300, 711299772)])
def test_gpu_build_random_networks(n_total, seed):
    print(f"gpu build compare N={n_total} seed={seed}", flush=True)
    rc, st = S.build_compare(n_total, seed, gpu=True)
    assert rc == 0, (rc, st)


def test_network_setup_gpu_through_c_abi():
    host = A.Network(30011, 5, setup="host")
    gpu = A.Network(30011, 5, setup="gpu-only")
    assert gpu.built_on_gpu and not host.built_on_gpu
    assert gpu.n_edges == host.n_edges
    for which in ("exc", "inh", "sic", "e2a"):
        for a, b in zip(host.csr(which), gpu.csr(which)):
            assert np.array_equal(a, b), which
    for p in A.Network.PARAM_NAMES:
        assert np.array_equal(host.params(p), gpu.params(p)), p


def test_gpu_only_declines_small_networks():
    with pytest.raises(Exception):
        A.Network(100, 1, setup="gpu-only")
    assert not A.Network(100, 1, setup="gpu").built_on_gpu
    # the automatic choice keeps mid-size networks on the host, large ones on the GPU
    assert not A.Network(20000, 1, setup="gpu").built_on_gpu
    assert A.Network(100000, 1, setup="gpu").built_on_gpu


@pytest.mark.parametrize("method,precision", [("plain", "fp64"), ("blocked", "fp64"), ("blocked", "fp32c")])
def test_device_resident_network_simulates_like_host_built(method, precision):
    # the GPU-built network stays in device memory and the engine copies it
    # device to device; the simulation must equal one of the host-built network
    print(f"device-resident network, {method} {precision}", flush=True)
    host = A.Network(24000, 17, setup="host")
    gpu = A.Network(24000, 17, setup="gpu-only")
    a = A.Simulation(host, method=method, precision=precision)
    b = A.Simulation(gpu, method=method, precision=precision)
    a.step(400)
    b.step(400)
    assert a.spike_count() == b.spike_count() > 0
    for f in ("V", "Ca"):
        assert np.array_equal(a.state(f), b.state(f)), f
    # the host copy of the connectivity, made on first request, is the same too
    for which in ("exc", "sic"):
        for x, y in zip(host.csr(which), gpu.csr(which)):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("method", ["plain", "blocked"])
def test_cli_summary_same_with_gpu_and_host_setup(method):
    # the drop-in CLI generates the network on the GPU by default: its
    # summary must equal the one from the host walk (timing lines aside)
    args = ["--n", "150000", "--seed", "9", "--pre-ms", "20", "--sim-ms", "60", "--method", method]
    timing = ("wall time", "steps/s")
    out = {}
    for setup in ("gpu", "host"):
        r = A.run_cli(args + ["--setup", setup], timeout=300)
        assert r.returncode == 0, r.stderr
        out[setup] = [l.split(", built")[0] for l in r.stdout.splitlines() if not l.startswith(timing)]
    assert out["gpu"] == out["host"]
