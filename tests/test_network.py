"""Network construction: the product's C++ setup against the oracle's C
restatement (bit-exact), plus invariants that do not depend on either
(reference: src/astrosim_fused.cu:595-735, 917-933)."""
import math

import numpy as np
import pytest

import astrosimgpu_cuda_amd as A
import pyoracle as O

# 300,001 cells build on several host threads (one per 65,536 cells), so the
# chunked inversion and the skip windows run in parallel there.
CASES = [(1, 1), (2, 3), (7, 1), (20, 5), (99, 2), (100, 1), (101, 4), (500, 1), (5000, 7), (40000, 11),
         (300001, 13)]


@pytest.mark.parametrize("n,seed", CASES)
def test_network_matches_oracle_bitwise(n, seed):
    a, o = A.Network(n, seed), O.OracleNet(n, seed)
    assert (a.n_exc, a.n_inh, a.n_astro) == (o.n_exc, o.n_inh, o.n_astro)
    assert a.p_primary == o.p_primary
    for w in ("exc", "inh", "sic", "e2a"):
        r1, s1 = a.csr(w)
        r2, s2 = o.csr(w)
        assert np.array_equal(r1, r2), w
        assert np.array_equal(s1, s2), w
    for w in ("Ca_tot", "IP3_0", "tau_IP3", "delta_IP3", "V_reset", "b"):
        assert np.array_equal(a.params(w).view(np.uint64), o.params(w).view(np.uint64)), w


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 9, 10, 500, 1000, 123457, 10_000_000])
def test_scaling_law_closed_form(n):
    # N_E = round(0.8N) (>=1), N_I = N - N_E (>=1), N_A = round(N/5) (>=1), p = min(1, 80/N_E)
    ne = max(1, int(np.floor(0.8 * n + 0.5)))
    ni = max(1, n - ne)
    na = max(1, int(np.floor(n / 5 + 0.5)))
    o = O.OracleNet(n, 1) if n <= 200000 else None
    if o is not None:
        assert (o.n_exc, o.n_inh, o.n_astro) == (ne, ni, na)
        assert o.p_primary == min(1.0, 80.0 / ne)


@pytest.fixture(scope="module")
def net20k():
    return A.Network(20000, 3)


def test_no_autapses_and_sorted_rows(net20k):
    n = net20k
    rs, src = n.csr("exc")
    tgt = np.repeat(np.arange(n.n_neurons), np.diff(rs))
    assert not np.any(src == tgt)
    assert np.all(src < n.n_exc)
    # rows list sources ascending and without repeats (Bernoulli per pair)
    same_row = tgt[1:] == tgt[:-1]
    assert np.all(src[1:][same_row] > src[:-1][same_row])
    rs, src = n.csr("inh")
    tgt = np.repeat(np.arange(n.n_neurons), np.diff(rs))
    assert np.all(src < n.n_inh)  # local inhibitory ids
    assert not np.any(src + n.n_exc == tgt)
    same_row = tgt[1:] == tgt[:-1]
    assert np.all(src[1:][same_row] > src[:-1][same_row])


def test_in_degree_statistics(net20k):
    n = net20k
    for w, k in (("exc", 80.0 * (n.n_exc - 1) / n.n_exc), ("inh", 80.0 * n.n_inh / n.n_exc)):
        rs, _ = n.csr(w)
        deg = np.diff(rs).astype(float)
        assert abs(deg.mean() - k) < 0.5, (w, deg.mean())
        # Binomial spread around K_SYN
        assert abs(deg.std() - np.sqrt(k)) < 0.1 * np.sqrt(k)


def test_third_factor_edges(net20k):
    n = net20k
    ne, ns = n.n_edges["e2a"], n.n_edges["sic"]
    assert ne == ns
    frac = ns / n.n_edges["exc"]
    assert abs(frac - 0.2) < 0.01
    rs, src = n.csr("sic")
    assert np.all(src < n.n_astro)
    # each neuron's astrocyte sources come from its pool of at most 5
    for j in range(0, n.n_neurons, 97):
        assert len(set(src[rs[j]:rs[j + 1]].tolist())) <= 5
    # astrocyte out-degree (count of n2a edges into it) within the cap
    rs2, src2 = n.csr("e2a")
    assert np.diff(rs2).max() <= 10000
    assert np.all(src2 < n.n_exc)
    # n2a and a2n are the two halves of the same third-factor synapses:
    # per astrocyte, its n2a in-degree equals its a2n out-degree
    outdeg = np.bincount(src, minlength=n.n_astro)
    assert np.array_equal(outdeg, np.diff(rs2))


def test_params_within_bounds(net20k):
    n = net20k
    bounds = {"Ca_tot": (1.8958264782990153, 0.9, 1.05), "IP3_0": (0.010925583871830699, 0.9, 1.05),
              "tau_IP3": (1058.7589568400845, 0.9, 1.05), "delta_IP3": (0.04911592657233997, 0.9, 1.05)}
    for w, (m, lo, hi) in bounds.items():
        v = n.params(w)
        assert v.min() >= m * lo and v.max() <= m * hi
        # normal_redraw is a truncated normal N(1, 0.1) on [lo, hi] (in units
        # of the mean); its mean has a closed form
        a, b = (lo - 1.0) / 0.1, (hi - 1.0) / 0.1
        phi = lambda x: np.exp(-x * x / 2) / np.sqrt(2 * np.pi)
        Phi = lambda x: 0.5 * (1 + math.erf(x / np.sqrt(2)))
        want = 1.0 + 0.1 * (phi(a) - phi(b)) / (Phi(b) - Phi(a))
        assert abs(v.mean() / m - want) < 0.004
    vr = n.params("V_reset")
    assert np.all(vr[: n.n_exc] <= -50.0 * 0.9) and np.all(vr[: n.n_exc] >= -50.0 * 1.1)
    assert np.all(vr[n.n_exc:] <= -60.0 * 0.9) and np.all(vr[n.n_exc:] >= -60.0 * 1.1)


def test_dump_edges_csv_matches_csr(tmp_path):
    n = A.Network(300, 2)
    p = tmp_path / "edges.csv"
    n.dump_edges(p)
    lines = p.read_text().splitlines()
    assert lines[0] == "type,source,target"
    total = sum(n.n_edges.values())
    assert len(lines) == 1 + total
    rs, src = n.csr("inh")
    first_inh = lines[1 + n.n_edges["exc"]]
    t0 = int(np.nonzero(np.diff(rs))[0][0])
    assert first_inh == f"inh_primary,{int(src[0]) + n.n_exc},{t0}"
    assert any(l.startswith("n2a,") and ",A" in l for l in lines)
    assert any(l.startswith("a2n,A") for l in lines)


def test_invalid_sizes():
    with pytest.raises(A.AstrosimError):
        A.Network(0, 1)
    with pytest.raises(A.AstrosimError):
        A.Network(2**31, 1)
