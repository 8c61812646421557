"""The parallel connectivity generator (csrc/as_gpubuild.h, the method of
the GPU network build) run over its host backend -- the same walk, chain,
path, repair and emit bodies, with loops for kernels and a stable sort for
the radix sort -- against the sequential host builder: all four CSRs must be
identical (src/astrosim_fused.cu:631-735). The sizes cover the regimes the
method distinguishes: a handful of pools per astrocyte at the smallest, and
the autapse repairs (tens per network at any size) throughout."""
import pytest

import selftest_math as S


@pytest.mark.parametrize("n_total,seed", [(2000, 1), (3001, 9), (5000, 7), (20000, 3), (50000, 11)])
def test_emulated_parallel_build_equals_sequential(n_total, seed):
    rc, st = S.build_compare(n_total, seed, gpu=False)
    assert rc == 0, (rc, st)
    # the speculation had work to do: repairs happened and were composed
    assert st["repairs0"] > 0 and st["records0"] >= st["repairs0"]
    assert st["max_shift"] <= 12


def test_small_networks_are_declined():
    # p_primary = 1 (no geometric skips) or too few astrocytes for rejection pools
    rc, _ = S.build_compare(100, 1, gpu=False)
    assert rc == 2


@pytest.mark.parametrize("n_total,seed", [(4096, 6), (33333, 7)])
def test_chains_walking_again_past_their_stored_starts(n_total, seed):
    # with 8 stored starts per chain almost every chain on the path is
    # stitched partly from a second walk from where it stood at its ninth
    rc, st = S.build_compare(n_total, seed, gpu=False, cap_b=8)
    assert rc == 0, (rc, st)
