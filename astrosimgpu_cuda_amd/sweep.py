"""Size sweep over the drop-in CLI, as the reference's GH200 batch script.

    python -m astrosimgpu_cuda_amd.sweep [--sizes "500 5000 100000 1000000"]
        [--seed 1] [--pre-ms 1000] [--sim-ms 5000] [--bin PATH] [--csv FILE]
        [-- extra CLI flags, e.g. --method blocked --precision fp32c]

Runs the binary once per N_total with --n/--seed/--pre-ms/--sim-ms, echoes
its output indented, and prints one row per size with wall seconds, us/step,
recorded spikes and mean rate (scripts/run_gh200.sbatch:48-62 of the
reference; same defaults and column order). The reference's awk on the wall
line prints the literal "s" instead of the seconds (its field 4 of
"wall time : X s for ..."); this harness parses the number.
"""
from __future__ import annotations

import argparse
import csv
import re
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent

_WALL = re.compile(r"^wall time\s*:\s*([0-9.]+) s for (\d+) steps \(([0-9.]+) us/step")
_SPIKES = re.compile(r"^spikes \(recorded win\):\s*(\d+)")
_RATE = re.compile(r"^mean firing rate\s*:\s*([0-9.]+) Hz")
_STEPS_S = re.compile(r"^steps/s\s*:\s*([0-9.]+)")
_CA = re.compile(r"^mean calcium \(final\)\s*:\s*([0-9.eE+-]+)")
_ACTIVE = re.compile(r"^astrocytes active\s*:\s*(\d+) / (\d+)")
_ENGAGED = re.compile(r"^astrocytes engaged\s*:\s*(\d+) / (\d+)")


def parse_summary(text: str) -> dict:
    """The summary lines of one run (src/astrosim_fused.cu:1131-1141 of the
    reference); keys are missing for lines that are absent (a failed run)."""
    out: dict = {}
    for line in text.splitlines():
        line = line.strip()
        if m := _WALL.match(line):
            out["wall_s"], out["steps"], out["us_per_step"] = float(m[1]), int(m[2]), float(m[3])
        elif m := _SPIKES.match(line):
            out["spikes"] = int(m[1])
        elif m := _RATE.match(line):
            out["mean_hz"] = float(m[1])
        elif m := _STEPS_S.match(line):
            out["steps_per_s"] = float(m[1])
        elif m := _CA.match(line):
            out["mean_ca"] = float(m[1])
        elif m := _ACTIVE.match(line):
            out["active"], out["n_astro"] = int(m[1]), int(m[2])
        elif m := _ENGAGED.match(line):
            out["engaged"] = int(m[1])
    return out


def run_one(binary: str, n: int, seed: int, pre_ms: float, sim_ms: float, extra: list[str],
            timeout: float | None = None) -> tuple[str, dict, int]:
    cmd = [binary, "--n", str(n), "--seed", str(seed), "--pre-ms", f"{pre_ms:g}", "--sim-ms", f"{sim_ms:g}", *extra]
    p = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout)
    text = p.stdout + p.stderr
    return text, parse_summary(text), p.returncode


def sweep(sizes: list[int], seed: int = 1, pre_ms: float = 1000.0, sim_ms: float = 5000.0,
          binary: str | None = None, extra: list[str] | None = None, echo=print) -> list[dict]:
    binary = binary or str(PKG / "bin" / "astrosim_fused")
    if not Path(binary).exists():
        raise FileNotFoundError(f"{binary} is missing: build it with python -m astrosimgpu_cuda_amd.build")
    rows = []
    echo(f"{'N_total':<12} {'wall_s':>14} {'us/step':>14} {'spikes':>14} {'mean_Hz':>14}")
    for n in sizes:
        text, s, rc = run_one(binary, n, seed, pre_ms, sim_ms, extra or [])
        for line in text.splitlines():
            echo("  " + line)
        row = {"n_total": n, "returncode": rc, **s}
        rows.append(row)
        ok = rc == 0 and "wall_s" in s
        echo(f"{n:<12} {s['wall_s'] if ok else 'FAILED':>14} {s.get('us_per_step', '?'):>14} "
             f"{s.get('spikes', '?'):>14} {s.get('mean_hz', '?'):>14}")
    return rows


def main(argv: list[str] | None = None) -> int:
    argv = list(sys.argv[1:] if argv is None else argv)
    extra: list[str] = []
    if "--" in argv:
        i = argv.index("--")
        argv, extra = argv[:i], argv[i + 1:]
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--sizes", default="500 5000 100000 1000000")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--pre-ms", type=float, default=1000.0)
    ap.add_argument("--sim-ms", type=float, default=5000.0)
    ap.add_argument("--bin", default=None, help="the CLI to run (default: the built astrosim_fused)")
    ap.add_argument("--csv", default=None, help="also write the rows to this CSV file")
    a = ap.parse_args(argv)
    rows = sweep([int(x) for x in a.sizes.split()], a.seed, a.pre_ms, a.sim_ms, a.bin, extra)
    if a.csv:
        keys = ["n_total", "returncode", "wall_s", "steps", "us_per_step", "steps_per_s", "spikes", "mean_hz",
                "mean_ca", "active", "engaged", "n_astro"]
        with open(a.csv, "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=keys, extrasaction="ignore")
            w.writeheader()
            w.writerows(rows)
    return 0 if all(r["returncode"] == 0 and "wall_s" in r for r in rows) else 1


if __name__ == "__main__":
    sys.exit(main())
