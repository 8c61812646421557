"""The allocation-free spike read (as_sim_read_spikes_noalloc) against the
allocating one (as_sim_read_spikes), through the C ABI: two calls in a row,
the second on a smaller input, equal event lists; and its size limit."""
import numpy as np
import pytest

import astrosimgpu_cuda_amd as A

pytestmark = pytest.mark.gpu


def both(sim, first, n, cap=None):
    a = sim.spikes(first, n, cap=cap)
    b = sim.spikes(first, n, cap=cap, no_alloc=True)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    return a


def test_noalloc_equals_allocating_read():
    big = A.Simulation(A.Network(200_000, 3, setup="gpu"), method="plain", w_n2a=100.0, w_a2n=5000.0)
    big.step(1600)
    s, _ = both(big, 1590, 10)
    assert len(s) > 1000
    both(big, 1590, 10, cap=100)  # truncated to the cap
    small = A.Simulation(A.Network(5000, 9), method="blocked")
    small.step(600)
    s2, _ = both(small, 593, 3)
    assert len(s2) < len(s)
    both(big, 1595, 5)  # the big one again after the small one


def test_noalloc_rejects_caps_beyond_its_staging():
    sim = A.Simulation(A.Network(500, 1), method="plain")
    sim.step(10)
    with pytest.raises(A.AstrosimError, match="allocation-free"):
        sim.spikes(0, 10, cap=(1 << 20) + 1, no_alloc=True)
    s, _ = sim.spikes(0, 10, cap=1 << 20, no_alloc=True)
    assert len(s) == sim.spike_count()
