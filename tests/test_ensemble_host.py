"""Host side of the ensemble runner: the network-size draw is the counter
RNG's uniform (checked against the oracle's) mapped log-uniformly."""
import math

import numpy as np

import pyoracle as O
from astrosimgpu_cuda_amd import ensemble as E


def test_python_uniform_matches_oracle():
    for s, st, c in [(3, 0, 0), (3, 0, 255), (1000, 7, 12345), (2**63 + 5, 2**40, 3)]:
        assert E.splitmix_uniform(s, st, c) == O.rng_uniform(s, st, c)


def test_ensemble_sizes_log_uniform():
    sizes = E.ensemble_sizes(256)
    assert len(sizes) == 256
    assert min(sizes) >= 500 and max(sizes) <= 200000
    u = [(math.log(n) - math.log(500)) / (math.log(200000) - math.log(500)) for n in sizes]
    assert abs(np.mean(u) - 0.5) < 0.06
    assert sizes == E.ensemble_sizes(256)  # deterministic
    assert sizes[0] == int(round(math.exp(math.log(500) + O.rng_uniform(3, 0, 0) * math.log(400))))
