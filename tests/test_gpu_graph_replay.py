"""Captured HIP graphs of a count replay to the eager result, every time.

Each replay is checked against the eager count, with other eager work and
synchronisations between replays; tuning drops the graph."""
import numpy as np
import pytest

from test_gpu_parity import case, oracle_counts

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["rmat15", "planted", "karate", "star", "stencil3d"])
def test_plan_replay_equals_eager(gpu, orc, name):
    s, d, n = case(name, gpu)
    g = gpu.Graph.from_coo(s, d, n)
    want, want_pv = oracle_counts(orc, g)
    p = gpu.Plan(g)
    other = gpu.Plan(gpu.Graph.from_coo(*case("messy", gpu)))
    assert p.count()[0] == want
    p.capture()
    for _ in range(5):
        t, pv, ms = p.count()
        assert t == want and pv is None and ms > 0
        assert p.last_timing() == (-1.0, -1.0)
        t2, pv2, _ = p.count(per_vertex=True)  # per-vertex counts run eagerly
        assert t2 == want
        np.testing.assert_array_equal(pv2, want_pv)
        other.count()
    p.tune(group_lanes=16)  # drops the graph
    assert p.count()[0] == want and p.last_timing()[1] >= 0
    p.capture()
    assert all(p.count()[0] == want for _ in range(3))


def test_batch_replay_equals_eager(gpu):
    work = gpu.batch_workload(48, seed=7, n_lo=256, n_hi=4096)
    graphs = [gpu.Graph.from_coo(s, d, n) for s, d, n, _ in work]
    b = gpu.Batch(graphs)
    t0, per0, _ = b.count()
    single = [gpu.Plan(x).count()[0] for x in graphs[:8]]
    assert list(per0[:8]) == single
    b.capture()
    for _ in range(4):
        t, per, ms = b.count()
        assert t == t0 and ms > 0
        np.testing.assert_array_equal(per, per0)
        gpu.Plan(graphs[3]).count()
    b.tune(group_lanes=8, unroll=2)
    t, per, _ = b.count()
    assert t == t0
    np.testing.assert_array_equal(per, per0)


def test_overlapped_dense_and_wedge_kernels(gpu, orc):
    """A plan large enough to run the dense-tile kernel on a second stream
    beside the wedge kernel: eager, per-vertex and replayed counts agree
    with the oracle."""
    s, d = gpu.gen_rmat(18, 16, seed=18)
    g = gpu.Graph.from_coo(s, d, 1 << 18)
    want, want_pv = oracle_counts(orc, g)
    p = gpu.Plan(g, tiny_items=1)
    info = p.info
    assert info["dense_tiles"] > 0 and info["wedge_checks"] >= 1 << 24 and info["tiny_items"] > 0
    for _ in range(3):
        assert p.count()[0] == want
        t, pv, _ = p.count(per_vertex=True)
        assert t == want
        np.testing.assert_array_equal(pv, want_pv)
    p.capture()
    for _ in range(3):
        assert p.count()[0] == want
        p.count(per_vertex=True)


@pytest.mark.parametrize("name", ["rmat15", "planted", "star"])
def test_shares_add_up(gpu, orc, name):
    """The parts of a k-way split of a plan's work add up to its count."""
    s, d, n = case(name, gpu)
    g = gpu.Graph.from_coo(s, d, n)
    want, _ = oracle_counts(orc, g)
    p = gpu.Plan(g, tile_k=64, tau=0 if name == "planted" else -1)
    for k in (1, 2, 3, 8):
        assert sum(p.count_share(k, j)[0] for j in range(k)) == want, k
    with pytest.raises(gpu.BtcError):
        p.count_share(4, 4)
    assert p.count()[0] == want
