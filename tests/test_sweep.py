"""The size-sweep harness (the reference's scripts/run_gh200.sbatch): summary
parsing and the table, driven here by the host oracle's CLI, which prints the
same summary lines as astrosim_fused (src/astrosim_fused.cu:1131-1141)."""
import csv
from pathlib import Path

import pytest

from astrosimgpu_cuda_amd import sweep as S

ROOT = Path(__file__).resolve().parents[1]
ORACLE_CLI = ROOT / "oracle" / "build" / "astrosim_oracle"

SAMPLE = """=== done ===
wall time            : 0.123 s for 60000 steps (2.05 us/step, both kernels)
spikes (recorded win): 4321
mean firing rate     : 1.7284 Hz
mean calcium (final) : 0.0713254 uM  (SIC_th=0.19669 uM)
astrocytes active    : 3 / 100 (Ca above SIC threshold at the final step only)
astrocytes engaged   : 7 / 100 (Ca crossed SIC threshold at ANY point in the recorded window -- x)
steps/s              : 487804.878
"""


def test_parse_summary_fields():
    s = S.parse_summary(SAMPLE)
    assert s == {"wall_s": 0.123, "steps": 60000, "us_per_step": 2.05, "spikes": 4321, "mean_hz": 1.7284,
                 "mean_ca": 0.0713254, "active": 3, "n_astro": 100, "engaged": 7, "steps_per_s": 487804.878}


def test_parse_summary_of_a_failed_run_is_empty():
    assert S.parse_summary("unknown argument: --x (try --help)\n") == {}


@pytest.mark.skipif(not ORACLE_CLI.exists(), reason="oracle not built")
def test_sweep_over_the_oracle_cli(tmp_path):
    lines = []
    rows = S.sweep([60, 150], seed=3, pre_ms=5, sim_ms=10, binary=str(ORACLE_CLI), extra=["--threads", "1"],
                   echo=lines.append)
    assert [r["n_total"] for r in rows] == [60, 150]
    for r in rows:
        assert r["returncode"] == 0 and r["steps"] == 150
        assert r["spikes"] >= 0 and r["mean_hz"] >= 0.0
    assert lines[0].split() == ["N_total", "wall_s", "us/step", "spikes", "mean_Hz"]
    table = [ln for ln in lines if not ln.startswith("  ")][1:]
    assert [int(t.split()[0]) for t in table] == [60, 150]
    assert int(table[1].split()[3]) == rows[1]["spikes"]
    out = tmp_path / "sweep.csv"
    assert S.main(["--sizes", "60", "--seed", "3", "--pre-ms", "5", "--sim-ms", "10", "--bin", str(ORACLE_CLI),
                   "--csv", str(out), "--", "--threads", "1"]) == 0
    got = list(csv.DictReader(open(out)))
    assert got[0]["n_total"] == "60" and int(got[0]["steps"]) == 150
