"""bench.py's host-side contract: the multi-GPU refusal and the algorithmic
byte counts its roofline is computed from (the reference layout's bytes per
step, B(N, s) = N(16s + 20) + N_A(12s + 4) + 4 E, and the blocked kernel's own
state bytes)."""
import subprocess
import sys
from pathlib import Path
from types import SimpleNamespace

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
import bench  # noqa: E402


def fake_net(n, na, edges):
    return SimpleNamespace(n_neurons=n, n_astro=na, n_edges=edges)


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr
    assert p.stdout == ""


def test_reference_bytes_per_step_closed_form():
    # config 3 sizes with the expected edge counts 80N / 20N / 16N / 16N
    n, na = 10_000_000, 2_000_000
    net = fake_net(n, na, {"exc": 80 * n, "inh": 20 * n, "e2a": 16 * n, "sic": 16 * n})
    for prec, s in (("fp64", 8), ("fp32c", 4)):
        want = n * (16 * s + 20) + na * (12 * s + 4) + 4 * 132 * n + (24 * na if prec == "fp32c" else 0)
        assert bench.reference_bytes_per_step(net, prec) == want
    # 6.96 GB per step at FP64 (the figure the per-step HBM ceiling is quoted on)
    assert abs(bench.reference_bytes_per_step(net, "fp64") / 1e9 - 6.96) < 0.005


def test_blocked_kernel_state_bytes():
    net = fake_net(1000, 200, {"exc": 0, "inh": 0, "e2a": 0, "sic": 0})
    assert bench.kernel_bytes_per_launch(net, "fp64", "blocked") == 1000 * (16 * 8 + 36) + 200 * (10 * 8 + 13)
    assert bench.kernel_bytes_per_launch(net, "fp32c", "blocked") == 1000 * (16 * 4 + 36) + 200 * (16 * 4 + 13)
    # the plain kernel's algorithmic bytes are the reference layout's
    assert bench.kernel_bytes_per_launch(net, "fp64", "plain") == bench.reference_bytes_per_step(net, "fp64")


def test_roofline_uses_the_kernels_own_bytes_and_full_launch_time():
    n, na = 10_000_000, 2_000_000
    net = fake_net(n, na, {"exc": 80 * n, "inh": 20 * n, "e2a": 16 * n, "sic": 16 * n})
    leg = {"precision": "fp32c", "method": "blocked", "kernel_ms_per_full_launch": 1.0}
    r = bench.add_roofline(net, leg, "nonexistent_workload")
    own = n * (16 * 4 + 36) + na * (16 * 4 + 13)
    assert r["algorithmic_bytes_per_launch"] == own
    assert abs(r["achieved"] - own / 1e-3 / 1e9) < 1e-6
    assert abs(r["frac"] - r["achieved"] / 8000.0) < 1e-12
    assert r["frac"] < 1.0 and r["traffic"] is None
    plain = bench.add_roofline(net, {"precision": "fp64", "method": "plain", "kernel_ms_per_full_launch": 1.4},
                               "nonexistent_workload")
    assert plain["algorithmic_bytes_per_launch"] == bench.reference_bytes_per_step(net, "fp64")
