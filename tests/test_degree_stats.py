"""--dump-edges and its degree statistics (src/astrosim_fused.cu:873-902,
scripts/degree_stats.py:17-49), on the host: the library's dump of a network
equals, row for row, a dump written from the FP64 oracle's own in-CSRs (an
independent generator), the statistics read from the file equal those
computed from the oracle's CSRs, and the statistics' definitions (population
standard deviation, upper median) hold on a hand-made file."""
import numpy as np

import astrosimgpu_cuda_amd as A
import pyoracle as O
from astrosimgpu_cuda_amd import degree_stats as D


def test_dump_equals_the_oracle_dump_and_degrees_agree(tmp_path):
    n, seed = 3000, 4
    path = tmp_path / "edges.csv"
    A.Network(n, seed).dump_edges(path)
    onet = O.OracleNet(n, seed)
    got = path.read_text().splitlines()
    want = list(D.dump_rows(onet))
    assert len(got) == len(want) == 1 + sum(len(onet.csr(w)[1]) for w in ("exc", "inh", "e2a", "sic"))
    assert got == want
    assert D.csv_degrees(path) == D.csr_degrees(onet)
    # the dump's id spelling: inhibitory sources offset by N_E, astrocytes prefixed "A"
    inh = [l for l in got if l.startswith("inh_primary,")]
    assert min(int(l.split(",")[1]) for l in inh) >= onet.n_exc
    assert all(l.split(",")[2].startswith("A") for l in got if l.startswith("n2a,"))
    assert all(l.split(",")[1].startswith("A") for l in got if l.startswith("a2n,"))


def test_summary_definitions(tmp_path):
    p = tmp_path / "e.csv"
    p.write_text("type,source,target\n"
                 "exc_primary,0,1\nexc_primary,0,2\nexc_primary,0,3\nexc_primary,1,2\n"
                 "n2a,5,A0\n")
    rep = D.report(D.csv_degrees(p)).splitlines()
    # exc sources {0: 3, 1: 1}: mean 2, population std 1, upper median 3;
    # targets {1: 1, 2: 2, 3: 1}: mean 4/3, std sqrt(2/9), median 1
    assert rep[0] == "== exc_primary: 4 edges =="
    assert rep[1] == ("  out-degree (per source): n=      2  mean=   2.000  std=  1.000  "
                      "min=     1  median=     3  max=     3")
    assert rep[2] == ("  in-degree  (per target): n=      3  mean=   1.333  std=  0.471  "
                      "min=     1  median=     1  max=     2")
    assert rep[3] == "== n2a: 1 edges =="


def test_degree_means_follow_the_scaling_law():
    # fixed in-degree: exc in-rows ~ Binomial(N_E, 80/N_E) (mean 80, less the
    # autapse on excitatory rows), inh ~ Binomial(N_I, 80/N_E) = 20 at N_I = N_E/4
    onet = O.OracleNet(20000, 2)
    deg = D.csr_degrees(onet)
    exc_in = np.array(list(deg["exc_primary"][1].values()))
    inh_in = np.array(list(deg["inh_primary"][1].values()))
    assert abs(exc_in.mean() - 80.0) < 0.5
    assert abs(inh_in.mean() - 20.0) < 0.3
    # every n2a synapse has its a2n mirror: equal totals
    assert sum(deg["n2a"][0].values()) == sum(deg["a2n"][1].values())
