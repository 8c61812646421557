"""Batches of networks in one launch per block: each member's results are
bit-identical to running it alone (FP64 and FP32c), also when members join
late, leave, reuse freed ranges, when blocks are replayed from a captured
hipGraph, and when the autotuner changes the workgroup size."""
import numpy as np
import pytest

import astrosimgpu_cuda_amd as A

pytestmark = pytest.mark.gpu

FIELDS = ["Ca", "IP3", "h", "V", "w", "g_ex", "dg_ex", "g_in", "dg_in", "I_sic_hold", "stp_x", "stp_u",
          "stp_t_last"]


def alone(net, steps, precision="fp64"):
    sim = A.Simulation(net, method="blocked", precision=precision)
    sim.step(steps)
    out = {f: sim.state(f) for f in FIELDS}
    out["spikes"] = sim.spike_count()
    out["ever"] = sim.ever_active()
    sim.close()
    return out


def assert_member_equals(batch, m, ref, where):
    for f in FIELDS:
        g = batch.state(m, f)
        bad = np.nonzero(g.view(np.uint64) != ref[f].view(np.uint64))[0]
        assert bad.size == 0, f"{where} member {m} {f}: {bad.size} differ"
    assert batch.spike_count(m) == ref["spikes"], where
    assert np.array_equal(batch.ever_active(m), ref["ever"]), where


@pytest.fixture(scope="module")
def nets():
    return [A.Network(n, s) for n, s in ((500, 1), (2000, 2), (7000, 3), (1200, 4))]


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_members_match_alone(nets, precision):
    print(f"[batch] members vs alone {precision}", flush=True)
    steps = 300
    refs = [alone(n, steps, precision) for n in nets[:3]]
    b = A.Batch(nets[:3], precision=precision)
    ms = [b.add(n, steps) for n in nets[:3]]
    b.step(steps)
    for m, r in zip(ms, refs):
        assert_member_equals(b, m, r, "batch")
    b.close()


def test_late_join_leave_and_reuse(nets):
    print("[batch] staggered admission and range reuse", flush=True)
    cap = A.Batch.capacity_for(nets[:3], members=4)
    b = A.Batch(cap)
    m0 = b.add(nets[0], 500)
    m1 = b.add(nets[1], 200)
    b.step(100)                      # members join at block boundaries
    m2 = b.add(nets[2], 250)
    b.step(230)                      # m1 reaches its 200 steps inside this window and stops
    assert b.member_steps(m1) == (200, 200)
    ref1 = alone(nets[1], 200)
    assert_member_equals(b, m1, ref1, "finished member")
    b.remove(m1)
    b.step(20)                       # the freed ranges drain, then can be reused
    assert b.fits(nets[3])
    m3 = b.add(nets[3], 150)
    b.step(150)
    assert_member_equals(b, m0, alone(nets[0], 500), "m0")
    assert_member_equals(b, m2, alone(nets[2], 250), "m2 (joined at step 100)")
    assert_member_equals(b, m3, alone(nets[3], 150), "m3 (reused ranges)")
    b.close()


def test_graph_replays_match_eager(nets):
    print("[batch] graph replay vs eager", flush=True)
    g = A.Batch(nets[:3])
    e = A.Batch(nets[:3])
    gm = [g.add(n, 10**6) for n in nets[:3]]
    em = [e.add(n, 10**6) for n in nets[:3]]
    for rep in range(4):
        g.run_graph(5)               # 5 blocks = 50 steps per replay
        e.step(50)
        g.sync()
        for a, b_ in zip(gm, em):    # eager reads between replays
            for f in ("V", "Ca", "w"):
                assert np.array_equal(g.state(a, f).view(np.uint64), e.state(b_, f).view(np.uint64)), (rep, f)
            assert g.spike_count(a) == e.spike_count(b_), rep
    g.close()
    e.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_chunked_graph_replays_match(nets, precision):
    # 1024-thread tiles stage fewer steps per launch in LDS, so each block
    # is several integrate launches (also inside the captured graph); the
    # results do not depend on that split.
    print(f"[batch] chunked graph replay {precision}", flush=True)
    g = A.Batch(nets[:3], precision=precision, block=1024)
    e = A.Batch(nets[:3], precision=precision)
    gm = [g.add(n, 10**6) for n in nets[:3]]
    em = [e.add(n, 10**6) for n in nets[:3]]
    for rep in range(3):
        g.run_graph(5)
        e.step(50)
        g.sync()
        for a, b_ in zip(gm, em):
            for f in FIELDS:
                assert np.array_equal(g.state(a, f).view(np.uint64), e.state(b_, f).view(np.uint64)), (rep, f)
            assert g.spike_count(a) == e.spike_count(b_), rep
    g.close()
    e.close()


def test_add_many_matches_alone(nets):
    # bulk admission: uploads back to back, one wait; members with their own step counts
    print("[batch] add_many fp32c", flush=True)
    steps = [300, 200, 300]
    refs = [alone(n, k, "fp32c") for n, k in zip(nets[:3], steps)]
    b = A.Batch(nets[:3], precision="fp32c")
    ms = b.add_many(nets[:3], steps)
    assert sorted(ms) == sorted(set(ms)) and all(m >= 0 for m in ms)
    b.step(300)
    for m, ref, k in zip(ms, refs, steps):
        assert b.member_steps(m) == (k, k)
        assert_member_equals(b, m, ref, f"add_many member {m}")
    b.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_autotune_keeps_results(nets, precision):
    # the tuner runs real blocks at each (workgroup size, LDS chunk) pair
    print(f"[batch] autotune {precision}", flush=True)
    b = A.Batch(nets[:3], precision=precision)
    ms = [b.add(n, 600) for n in nets[:3]]
    chosen = b.autotune((64, 128, 256, 512), blocks_each=3)
    info = b.info()
    assert chosen in (64, 128, 256, 512)
    assert 1 <= info["chunk_steps"] <= info["block_steps"]
    assert 0 < info["global_step"] <= 8 * 3 * info["block_steps"]
    b.step(600 - info["global_step"])
    for m, n in zip(ms, nets[:3]):
        assert_member_equals(b, m, alone(n, 600, precision), f"autotuned ({chosen}, chunk {info['chunk_steps']})")
    b.close()
