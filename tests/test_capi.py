"""The C ABI library loads, exports every symbol include/astrosim.h declares,
and behaves correctly on the paths that need no GPU."""
import re
import subprocess
from pathlib import Path

import pytest

import astrosimgpu_cuda_amd as A

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "astrosim.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(as_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    out = subprocess.run(["nm", "-D", "--defined-only", str(A.lib_path())], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (as_\w+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing


def test_binding_covers_header():
    bound = {name for name, _, _ in A.SIGNATURES}
    assert set(declared_symbols()) == bound


def test_library_loads_and_reports_version():
    lib = A.load()
    assert b"gfx950" in lib.as_version()


def test_library_has_gfx950_code_object():
    blob = A.lib_path().read_bytes()
    assert b"gfx950" in blob


def test_config_validation_without_gpu():
    net = A.Network(500, 1)
    with pytest.raises(A.AstrosimError) as e:
        A.Simulation.footprint(net, block=100)
    assert e.value.status == 1
    with pytest.raises(A.AstrosimError):
        A.Simulation.footprint(net, dt=0.0)
    assert A.Simulation.footprint(net) > 0


def test_cli_help_and_unknown_flag():
    r = A.run_cli(["--help"])
    assert r.returncode == 0 and "--n N_total" in r.stdout
    r = A.run_cli(["--bogus"])
    assert r.returncode == 1 and "unknown argument: --bogus (try --help)" in r.stderr


def test_cli_dump_edges_runs_without_gpu(tmp_path):
    p = tmp_path / "e.csv"
    r = A.run_cli(["--n", "500", "--seed", "1", "--dump-edges", str(p)])
    assert r.returncode == 0, r.stderr
    assert "connectivity: exc=40073 inh=9957 e2a=8019 sic=8019 edges" in r.stdout
    assert f"Wrote edge list to {p}" in r.stdout
    assert p.read_text().startswith("type,source,target\n")


def test_cli_dump_edges_unwritable():
    r = A.run_cli(["--n", "50", "--dump-edges", "/nonexistent_dir/x.csv"])
    assert r.returncode == 1
    assert "cannot open /nonexistent_dir/x.csv for --dump-edges" in r.stderr
