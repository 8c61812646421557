"""The fixed-divisor division of the FP64 step (as_model.h `divide`) against
IEEE division on the host, bit for bit: random dividends, signed zeros, the
simulator's own divisors, per-cell tau_IP3 divisors, and quotients within
2^-100 of a rounding midpoint (where a method that is only nearly correctly
rounded fails)."""
import numpy as np

import selftest_math as S


def _check(a, b):
    got = S.evaluate(S.DIVIDE, a, b)
    want = a / b
    bad = np.nonzero(S.bits(got) != S.bits(want))[0]
    assert bad.size == 0, [(float(a[i]), float(b[i]), float(got[i]), float(want[i])) for i in bad[:5]]


def test_simulator_divisors_random_dividends():
    rng = np.random.default_rng(11)
    for b in S.SIM_DIVISORS:
        a = S.random_dividends(rng, 100_000)
        _check(a, np.full_like(a, b))


def test_tau_ip3_divisors():
    # per-astrocyte tau_IP3 = 1058.76 scaled by the randomisation's [0.9, 1.05]
    rng = np.random.default_rng(12)
    b = 1058.7589568400845 * rng.uniform(0.9, 1.05, 200_000)
    a = rng.uniform(-1e-2, 1e-2, b.size)
    _check(a, b)


def test_signed_zero_quotients():
    b = np.array(S.SIM_DIVISORS * 2)
    a = np.array([0.0] * len(S.SIM_DIVISORS) + [-0.0] * len(S.SIM_DIVISORS))
    _check(a, b)


def test_near_midpoint_quotients():
    rng = np.random.default_rng(13)
    a, b = S.near_midpoint_quotients(rng, 100_000)
    _check(a, b)
    a, b = S.near_midpoint_quotients(rng, 40_000, divisors=S.SIM_DIVISORS)
    _check(a, b)


def test_generator_is_hard():
    # the cases really sit next to midpoints: RN(a * RN(1/b)) misses the
    # correctly rounded quotient on a visible share of them
    rng = np.random.default_rng(14)
    a, b = S.near_midpoint_quotients(rng, 20_000)
    naive = a * (1.0 / b)
    assert np.count_nonzero(S.bits(naive) != S.bits(a / b)) > 1000
