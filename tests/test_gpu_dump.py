"""--dump-edges of networks generated on the GPU (src/astrosim_fused.cu:
873-902) against a dump written from the FP64 oracle's own in-CSRs, row for
row, through the C ABI and through the drop-in CLI (--setup gpu)."""
import numpy as np
import pandas as pd
import pytest

import astrosimgpu_cuda_amd as A
import pyoracle as O
from astrosimgpu_cuda_amd import degree_stats as D

pytestmark = pytest.mark.gpu

SPEC = (("exc_primary", "exc", False, "", ""), ("inh_primary", "inh", True, "", ""),
        ("n2a", "e2a", False, "", "A"), ("a2n", "sic", False, "A", ""))


def assert_dump_equals_oracle(path, onet):
    df = pd.read_csv(path, dtype=str)
    assert list(df.columns) == ["type", "source", "target"]
    pos = 0
    for t, which, offset, sp, tp in SPEC:
        rs, src = onet.csr(which)
        n = len(src)
        part = df.iloc[pos:pos + n]
        assert (part["type"] == t).all(), t
        want_src = src.astype(np.int64) + (onet.n_exc if offset else 0)
        want_tgt = np.repeat(np.arange(len(rs) - 1), np.diff(rs.astype(np.int64)))
        got_src = part["source"].str.slice(len(sp)).astype(np.int64).to_numpy()
        got_tgt = part["target"].str.slice(len(tp)).astype(np.int64).to_numpy()
        if sp:
            assert part["source"].str.startswith(sp).all()
        if tp:
            assert part["target"].str.startswith(tp).all()
        assert np.array_equal(got_src, want_src), t
        assert np.array_equal(got_tgt, want_tgt), t
        pos += n
    assert pos == len(df)


def test_gpu_built_network_dump(tmp_path):
    net = A.Network(20000, 3, setup="gpu-only")
    assert net.built_on_gpu
    path = tmp_path / "edges.csv"
    net.dump_edges(path)
    onet = O.OracleNet(20000, 3)
    assert_dump_equals_oracle(path, onet)
    assert D.csv_degrees(path) == D.csr_degrees(onet)


def test_cli_dump_with_gpu_setup(tmp_path):
    path = tmp_path / "edges.csv"
    r = A.run_cli(["--n", "120000", "--seed", "8", "--setup", "gpu", "--dump-edges", str(path)], timeout=300)
    assert r.returncode == 0, r.stderr
    assert f"Wrote edge list to {path}" in r.stdout
    assert_dump_equals_oracle(path, O.OracleNet(120000, 8))
