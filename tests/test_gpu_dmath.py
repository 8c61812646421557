"""Device against host bits for the deterministic math of the FP64 path
(as_dmath.h exp/log/cos, as_model.h fixed-divisor division): every input of
wide random samples and of the edge ranges gives identical bits on the GPU
and on the host. The host side is checked against the C library and IEEE
division by test_dmath.py and test_divide.py."""
import math

import numpy as np
import pytest

import selftest_math as S

pytestmark = pytest.mark.gpu


def _same(fn, x, y=None):
    host = S.evaluate(fn, x, y)
    dev = S.evaluate(fn, x, y, device=True)
    bad = np.nonzero(S.bits(host) != S.bits(dev))[0]
    assert bad.size == 0, [(float(x[i]), float(host[i]), float(dev[i])) for i in bad[:5]]
    return dev


def test_exp_device_matches_host():
    rng = np.random.default_rng(21)
    x = np.concatenate([
        rng.uniform(-750.0, 712.0, 400_000),
        rng.uniform(-30.0, 30.0, 400_000),         # the AdEx spike-current range
        rng.uniform(-745.2, -708.0, 200_000),      # subnormal results
        rng.uniform(709.0, 709.782712893384, 10_000),
        [0.0, -0.0, 709.782712893384, 709.8, -745.2, -745.3, math.inf, -math.inf, math.nan, 1e-300, -1e-300],
    ])
    y = _same(S.EXP, x)
    assert y[-11] == 1.0 and math.isinf(y[-5]) and y[-4] == 0.0 and math.isnan(y[-3])


def test_log_device_matches_host():
    rng = np.random.default_rng(22)
    x = np.concatenate([
        np.exp(rng.uniform(-740.0, 709.0, 400_000)),
        rng.uniform(0.0, 1.0, 200_000),
        1.0 + rng.uniform(-1e-3, 1e-3, 200_000),
        [5e-324, 2.2250738585072014e-308, 1.0, 0.0, -1.0, math.inf, math.nan],
    ])
    _same(S.LOG, x)


def test_cos_device_matches_host():
    rng = np.random.default_rng(23)
    x = np.concatenate([
        6.283185307179586 * rng.uniform(0.0, 1.0, 400_000),  # the Box-Muller range
        rng.uniform(-2.0**20, 2.0**20, 200_000),
        [0.0, math.pi / 2, math.pi, 1.5 * math.pi, 2.0 ** 30, math.nan],
    ])
    _same(S.COS, x)


def test_fixed_divisor_device_matches_ieee():
    rng = np.random.default_rng(24)
    parts_a, parts_b = [], []
    for b in S.SIM_DIVISORS:
        a = S.random_dividends(rng, 50_000)
        parts_a.append(a)
        parts_b.append(np.full_like(a, b))
    a, b = S.near_midpoint_quotients(rng, 100_000)
    parts_a.append(a)
    parts_b.append(b)
    tau = 1058.7589568400845 * rng.uniform(0.9, 1.05, 100_000)
    parts_a.append(rng.uniform(-1e-2, 1e-2, tau.size))
    parts_b.append(tau)
    a, b = np.concatenate(parts_a), np.concatenate(parts_b)
    dev = _same(S.DIVIDE, a, b)
    assert np.array_equal(S.bits(dev), S.bits(a / b))
    # and the device's own IEEE division agrees with the host's
    assert np.array_equal(S.bits(S.evaluate(S.DIVIDE_IEEE, a, b, device=True)), S.bits(a / b))


def _stream_bases(rng, n):
    return rng.integers(0, 2**63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)


def test_normal_device_matches_host():
    z = _stream_bases(np.random.default_rng(25), 500_000)
    _same(S.NORMAL, z.view(np.float64))


def test_normal32_within_bound_of_fp64_draw():
    # FP32 state draws its Gaussian noise in FP32 from the same uniforms; its
    # difference from the FP64 draw is bounded by 1e-5 absolute + 1e-6
    # relative. The absolute part is the tails near cos = 0: rounding u2 to
    # float moves the angle by up to 3.7e-7 rad, times a radius up to ~6.
    # Measured on MI355X: max abs 6.8e-6, max rel (|x| > 1) 9.7e-7.
    z = _stream_bases(np.random.default_rng(26), 1_000_000).view(np.float64)
    d64 = S.evaluate(S.NORMAL, z)
    d32 = S.evaluate(S.NORMAL32, z, device=True)
    err = np.abs(d32 - d64)
    print(f"normal32 vs FP64 draw: max abs {err.max():.3g}, max rel (|x|>1) "
          f"{(err[np.abs(d64) > 1] / np.abs(d64[np.abs(d64) > 1])).max():.3g}, mean abs {err.mean():.3g}")
    assert np.all(err <= 1e-5 + 1e-6 * np.abs(d64))
    assert abs(d32.mean()) < 0.01 and abs(d32.std() - 1.0) < 0.01


def test_spike_current32_within_one_ulp():
    # The FP32 state's AdEx spike current g_L*Delta_T/C_m*exp((V - V_T)/Delta_T)
    # (exponent formed and split in FP64, 2^f by the hardware exponential)
    # against the same current in long double, at every 2^-10 mV step of
    # [-120, 0] mV (V_peak = 0) and at random FP32 V, both populations.
    # Bound: 1 ulp of the FP32 result (measured: at most 0.83 ulp, 94% correctly
    # rounded; the hardware exponential is within 1 ulp and the fraction's
    # rounding to FP32 adds at most 2^-26 relative).
    rng = np.random.default_rng(31)
    grid = np.arange(-120.0, 0.0 + 2.0 ** -11, 2.0 ** -10)
    v = np.concatenate([grid, rng.uniform(-120.0, 0.0, 200_000).astype(np.float32).astype(np.float64)])
    for pop in (0.0, 1.0):
        y = np.full_like(v, pop)
        dev = S.evaluate(S.SPIKE32, v, y, device=True)
        ref = S.evaluate(S.SPIKE_EXACT, v, y)
        ok = ref > np.finfo(np.float32).tiny  # normal FP32 results (deep below threshold the current underflows)
        ulp = np.spacing(np.abs(ref[ok]).astype(np.float32)).astype(np.float64)
        err = np.abs(dev[ok] - ref[ok]) / ulp
        exact = np.mean(dev[ok] == ref[ok].astype(np.float32).astype(np.float64))
        print(f"[spike32] population {int(pop)}: max {err.max():.3f} ulp, mean {err.mean():.3f} ulp, "
              f"{exact:.4f} correctly rounded, {np.count_nonzero(ok)} normal results", flush=True)
        assert err.max() < 1.0, err.max()
        # below the normal range the current rounds to zero or a subnormal, as FP32 does
        assert np.all(np.abs(dev[~ok]) <= np.finfo(np.float32).tiny)
