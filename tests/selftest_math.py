"""Loader for the test-only math library (lib/libastrosim_selftest.so) and
input generators for the deterministic-math tests."""
from __future__ import annotations

import ctypes as C
import math
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
LIB = ROOT / "astrosimgpu_cuda_amd" / "lib" / "libastrosim_selftest.so"

EXP, LOG, COS, DIVIDE, DIVIDE_IEEE, NORMAL, NORMAL32, SPIKE32, SPIKE_EXACT = 0, 1, 2, 3, 4, 5, 6, 7, 8

# The fixed divisors the FP64 step uses with the default biology
# (C_m, tau_w and Delta_T of both populations, ratio_ER_cyt, the STP
# recovery time constant, h_step = dt = 0.1 as the refractory divisor).
SIM_DIVISORS = [130.0, 104.0, 450.0, 264.0, 2.0, 0.185, 800.0, 0.1]

_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise FileNotFoundError(f"{LIB} is missing: build with python -m astrosimgpu_cuda_amd.build")
        lib = C.CDLL(str(LIB))
        for name in ("as_selftest_eval_host", "as_selftest_eval_device"):
            fn = getattr(lib, name)
            fn.restype = C.c_int
            fn.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]
        lib.as_selftest_build_compare.restype = C.c_int
        lib.as_selftest_build_compare.argtypes = [C.c_uint32, C.c_uint64, C.c_int, C.POINTER(C.c_double),
                                                  C.c_uint32]
        _lib = lib
    return _lib


def evaluate(fn: int, x, y=None, device: bool = False) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, dtype=np.float64)
    assert x.shape == y.shape
    out = np.empty_like(x)
    f = load().as_selftest_eval_device if device else load().as_selftest_eval_host
    rc = f(fn, x.ctypes.data, y.ctypes.data, out.ctypes.data, x.size)
    if rc != 0:
        raise RuntimeError(f"as_selftest_eval_{'device' if device else 'host'}({fn}) returned {rc}")
    return out


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def near_midpoint_quotients(rng: np.random.Generator, n: int, divisors=None):
    """Pairs (a, b) whose quotient lies within r * 2^-106 (relative, r <= 64)
    of a midpoint between two doubles: the hardest cases for correct rounding.

    With b = B * 2^e (B odd) and a 54-bit odd M, choose M = r * B^-1 mod 2^54,
    so that B*M - r is a multiple of 2^54; then a = (B*M - r) / 2^54 * 2^e is a
    double and a/b = M/2^54 - r/(2^54 B), where M/2^54 is a midpoint of the
    doubles in [1/2, 1)."""
    a_out, b_out = [], []
    while len(a_out) < n:
        if divisors is not None:
            b = float(divisors[int(rng.integers(len(divisors)))])
        else:
            b = math.ldexp(1.0 + float(rng.integers(0, 2**52)) * 2.0**-52, int(rng.integers(-12, 12)))
        m, e = math.frexp(b)
        B = int(m * 2**53)
        eb = e - 53
        tz = (B & -B).bit_length() - 1
        Bo = B >> tz
        r = int(rng.integers(1, 65)) * (1 if rng.integers(2) else -1)
        M = (r * pow(Bo, -1, 2**54)) % 2**54
        if M < 2**53 or M % 2 == 0:
            continue
        A, rem = divmod(Bo * M - r, 2**54)
        assert rem == 0
        if not 0 < A < 2**53:
            continue
        a = math.ldexp(A, eb + tz)
        a_out.append(-a if rng.integers(2) else a)
        b_out.append(b)
    return np.array(a_out), np.array(b_out)


def random_dividends(rng: np.random.Generator, n: int) -> np.ndarray:
    """Dividends of both signs over 2^-40..2^40, with signed zeros."""
    a = np.ldexp(1.0 + rng.random(n), rng.integers(-40, 40, n)) * np.where(rng.random(n) < 0.5, -1.0, 1.0)
    a[:4] = [0.0, -0.0, 0.0, -0.0]
    return a


BUILD_STATS = ["counters", "chains", "ambiguous", "host_rows0", "host_rows1", "host_rows2", "records0",
               "records2", "repairs0", "repairs2", "max_shift", "gen_seconds", "host_seconds", "edges", "diff"]


def build_compare(n_total: int, seed: int, gpu: bool, cap_b: int = 0) -> tuple[int, dict]:
    """The parallel connectivity generator (host emulation, or the GPU build)
    against the sequential host builder: (0 identical / 1 different / 2
    declined / -1 error, statistics). cap_b: row starts a chain stores past
    its region in the emulation (0: the default)."""
    st = (C.c_double * 32)()
    rc = load().as_selftest_build_compare(n_total, seed, 1 if gpu else 0, st, cap_b)
    d = dict(zip(BUILD_STATS, list(st)))
    d["marks"] = list(st)[16:27]
    d["pools_seconds"], d["params_wait_seconds"] = st[27], st[28]
    d["max_b_draws"], d["max_repair_rows"], d["params_seconds"] = st[29], st[30], st[31]
    return rc, d
