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


@pytest.mark.parametrize("n_total,seed", [(2000, 1), (20000, 3), (145000, 42), (1000000, 7)])
def test_gpu_build_equals_host_build(n_total, seed):
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
