"""Calcium-transient analysis of a recording (host only): onsets of the
reference's SIC activity test, (Ca - SIC_th)*1000 > 1
(src/astrosim_fused.cu:416-417, 1117), counted per astrocyte."""
import math

import numpy as np

from astrosimgpu_cuda_amd import recording as R

TH = R.SIC_TH


def test_transients_count_onsets_and_engaged():
    above, below = TH + 0.002, TH + 0.0005  # (Ca - th)*1000 = 2 > 1, 0.5 <= 1
    ca = np.array([
        [below, above, below, above],
        [above, above, below, above],
        [below, above, below, below],
        [above, below, below, above],
        [above, below, below, above],
    ])
    t = R.transients(ca)
    # astro 0: onsets at samples 1 and 3; astro 1: active from the start (1);
    # astro 2: never; astro 3: 0 (start) and 3
    assert t["per_astro"].tolist() == [2, 1, 0, 2]
    assert t["total"] == 5 and t["engaged"] == 3
    assert t["mean_length"] == (3 + 3 + 0 + 4) / 5


def test_exactly_at_threshold_is_inactive():
    # (Ca - SIC_th)*1000 must exceed 1: the boundary itself is not active
    ca = np.array([[TH + 0.001 - 1e-12], [TH + 0.001 + 1e-9]])
    assert R.transients(ca)["per_astro"].tolist() == [1]
    assert R.transients(ca[:1])["total"] == 0


def test_empty_record():
    t = R.transients(np.zeros((0, 4)))
    assert t["total"] == 0 and t["engaged"] == 0 and math.isnan(t["mean_length"])


def test_min_delay_steps():
    assert [R.min_delay_steps(dt) for dt in (0.1, 0.05, 0.3, 1.0, 12.0)] == [10, 20, 3, 1, 1]
