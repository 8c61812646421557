"""Deterministic exp/log/cos (as_dmath.h) against the C library: at most
1 ulp apart over wide random samples, exact special values."""
import math

import numpy as np
import pytest

import pyoracle as O


def ulps(a, b):
    ia = np.array([a], dtype=np.float64).view(np.int64)[0]
    ib = np.array([b], dtype=np.float64).view(np.int64)[0]
    return abs(int(ia) - int(ib))


def _check(fn, ref, xs, max_ulp=1):
    worst = max(ulps(fn(float(x)), ref(float(x))) for x in xs)
    assert worst <= max_ulp, worst


def test_exp_range():
    rng = np.random.default_rng(0)
    _check(O.load().orc_exp, math.exp, rng.uniform(-745, 709, 20000))
    _check(O.load().orc_exp, math.exp, rng.uniform(-30, 30, 20000))


def test_log_range():
    rng = np.random.default_rng(1)
    _check(O.load().orc_log, math.log, np.exp(rng.uniform(-700, 700, 20000)))
    _check(O.load().orc_log, math.log, rng.uniform(0, 1, 20000)[1:])
    _check(O.load().orc_log, math.log, 1 + rng.uniform(-1e-3, 1e-3, 20000))


def test_cos_box_muller_range():
    rng = np.random.default_rng(2)
    xs = 6.283185307179586 * rng.uniform(0, 1, 40000)
    lib = O.load()
    worst = max(ulps(lib.orc_cos(float(x)), math.cos(float(x))) for x in xs if abs(math.cos(x)) > 1e-3)
    assert worst <= 1
    # near the zeros compare absolutely
    for x in xs:
        assert abs(lib.orc_cos(float(x)) - math.cos(float(x))) < 2e-16


def test_special_values():
    lib = O.load()
    assert lib.orc_exp(0.0) == 1.0
    assert lib.orc_exp(-1e12) == 0.0
    assert math.isinf(lib.orc_exp(710.0))
    assert lib.orc_log(1.0) == 0.0
    assert lib.orc_log(2.0) == pytest.approx(math.log(2.0), rel=0, abs=2.3e-16)
    assert math.isinf(lib.orc_log(0.0)) and lib.orc_log(0.0) < 0
    assert math.isnan(lib.orc_log(-1.0))
    assert lib.orc_log(5e-324) == pytest.approx(math.log(5e-324), rel=1e-15)
    assert lib.orc_cos(0.0) == 1.0
    assert lib.orc_cos(math.pi) == -1.0
    # exact powers of two round trip
    for k in range(-1000, 1000, 37):
        assert lib.orc_log(2.0 ** k) == pytest.approx(k * math.log(2.0), rel=1e-15, abs=1e-300)
