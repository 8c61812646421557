"""The sweep harness over the GPU CLI: each size's recorded spikes and mean
rate equal the FP64 host oracle CLI's (both methods are bit-identical to it)."""
from pathlib import Path

import pytest

from astrosimgpu_cuda_amd import sweep as S

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("method", ["plain", "blocked"])
def test_sweep_matches_oracle_cli(method):
    print(f"[sweep] {method}", flush=True)
    kw = dict(seed=1, pre_ms=20, sim_ms=100, echo=lambda *_: None)
    gpu = S.sweep([500, 3000], extra=["--method", method], **kw)
    ref = S.sweep([500, 3000], binary=str(ROOT / "oracle" / "build" / "astrosim_oracle"), **kw)
    for g, o in zip(gpu, ref):
        assert g["returncode"] == 0, g
        assert (g["steps"], g["spikes"], g["mean_hz"]) == (o["steps"], o["spikes"], o["mean_hz"])
        assert (g["active"], g["engaged"], g["n_astro"]) == (o["active"], o["engaged"], o["n_astro"])
