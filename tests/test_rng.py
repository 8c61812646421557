"""Counter RNG of the oracle pinned by known answers, exact identities and
brute force (reference: src/astrosim_fused.cu:127-216)."""
import math

import numpy as np
import pytest

import pyoracle as O

# splitmix64 from state 0: the first outputs of the published generator
# (Vigna, splitmix64.c). rng_bits(0, k, 0) is its k-th output because the
# stream term adds exactly k golden-gamma increments.
SPLITMIX64_SEED0 = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_splitmix64_known_answers():
    for k, want in enumerate(SPLITMIX64_SEED0, start=1):
        assert O.rng_bits(0, k, 0) == want


def test_counter_term_is_second_weyl_increment():
    # counter c adds c*0xBF58476D1CE4E5B9 to the pre-mix state: equal
    # pre-mix states give equal outputs.
    g1, g2 = 0x9E3779B97F4A7C15, 0xBF58476D1CE4E5B9
    for seed, c in [(5, 1), (123, 7), (2**63, 3)]:
        seed2 = (seed + c * g2 - g1) % 2**64
        assert O.rng_bits(seed, 0, c) == O.rng_bits(seed2, 1, 0)


def test_uniform_is_top_53_bits():
    rng = np.random.default_rng(1)
    for _ in range(200):
        s, st, c = (int(x) for x in rng.integers(0, 2**63, 3))
        u = O.rng_uniform(s, st, c)
        assert u == (O.rng_bits(s, st, c) >> 11) * 2.0**-53
        assert 0.0 <= u < 1.0


def _uniforms(seed, stream, n):
    return [O.rng_uniform(seed, stream, c) for c in range(n)]


@pytest.mark.parametrize("lam", [0.27, 0.25, 3.03e-4, 2.0, 12.5])
def test_poisson_knuth_brute_force(lam):
    limit = math.exp(-lam)  # exp(-lam) differs from the oracle's by <= 1ulp; pick streams away from ties
    for stream in range(300):
        us = _uniforms(77, stream, 64)
        prod, count = us[0], 0
        while prod > limit:
            count += 1
            prod *= us[count]
        assert O.rng_poisson(77, stream, lam) == count


@pytest.mark.parametrize("lam", [0.27, 4.0])
def test_poisson_moments(lam):
    x = np.array([O.rng_poisson(9, s, lam) for s in range(40000)], dtype=float)
    se = math.sqrt(lam / x.size)
    assert abs(x.mean() - lam) < 5 * se
    assert abs(x.var() - lam) < 0.06 * lam + 5 * se


def test_poisson_normal_branch_and_zero():
    assert O.rng_poisson(1, 2, 0.0) == 0
    assert O.rng_poisson(1, 2, -1.0) == 0
    lam = 50.0
    for s in range(50):
        u1, u2 = _uniforms(3, s, 2)
        z = math.sqrt(-2.0 * math.log(max(u1, 1e-300))) * math.cos(2 * math.pi * u2)
        v = lam + math.sqrt(lam) * z
        want = 0 if v < 0 else int(v + 0.5)
        assert abs(O.rng_poisson(3, s, lam) - want) <= (1 if abs((v + 0.5) % 1) < 1e-9 else 0)


def test_normal_box_muller_identity():
    for s in range(500):
        u1, u2 = _uniforms(11, s, 2)
        want = math.sqrt(-2.0 * math.log(u1)) * math.cos(6.283185307179586 * u2)
        got = O.rng_normal(11, s)
        assert abs(got - want) <= 4e-16 * max(1.0, abs(want)) * 4


def test_normal_moments():
    z = np.array([O.rng_normal(5, s) for s in range(50000)])
    assert abs(z.mean()) < 0.03
    assert abs(z.std() - 1.0) < 0.02


def test_hostrng_pairs_share_radius():
    # HostRng.normal(): cos value first, then the cached sin value of the same
    # (u1, u2); consecutive pairs therefore satisfy c^2 + s^2 = -2 ln u1.
    seed, stream = 42, 0x5EED
    z = O.hostrng_normals(seed, stream, 200)
    for i in range(100):
        u1 = O.rng_uniform(seed, stream, 2 * i)
        u2 = O.rng_uniform(seed, stream, 2 * i + 1)
        r2 = -2.0 * math.log(u1)
        assert z[2 * i] ** 2 + z[2 * i + 1] ** 2 == pytest.approx(r2, rel=1e-13)
        theta = 6.283185307179586 * u2
        assert z[2 * i] == pytest.approx(math.sqrt(r2) * math.cos(theta), rel=1e-12, abs=1e-13)
        assert z[2 * i + 1] == pytest.approx(math.sqrt(r2) * math.sin(theta), rel=1e-12, abs=1e-13)


def test_normal_from_stream_base_matches_oracle():
    # the kernels draw Gaussians from a precomputed stream base
    # (seed + stream * golden gamma); the draw equals the oracle's rng_normal
    import selftest_math as S
    rng = np.random.default_rng(15)
    seeds = [int(s) for s in rng.integers(0, 2**63, 400)]
    streams = [int(s) for s in rng.integers(0, 2**63, 400)]
    z = np.array([(s + st * 0x9E3779B97F4A7C15) % 2**64 for s, st in zip(seeds, streams)], dtype=np.uint64)
    got = S.evaluate(S.NORMAL, z.view(np.float64))
    want = np.array([O.rng_normal(s, st) for s, st in zip(seeds, streams)])
    assert np.array_equal(S.bits(got), S.bits(want))
