"""Batched counting: every graph's count from the one persistent launch
equals its own oracle count, and equals the count of the same graph alone."""
import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu


def oracle_total(orc, g):
    rp, c = g.csr()
    return orc.tc_merge(rp, c, threads=4)


def test_batch_workload_matches_oracle(gpu, orc):
    work = gpu.batch_workload(count=96, seed=7, n_lo=64, n_hi=3000)
    gs = [gpu.Graph.from_coo(s, d, n) for s, d, n, _ in work]
    want = np.array([oracle_total(orc, g) for g in gs], np.uint64)
    b = gpu.Batch(gs)
    total, per, ms = b.count()
    np.testing.assert_array_equal(per, want)
    assert total == int(want.sum()) and ms > 0
    assert b.info["graphs"] == 96 and b.info["n"] % 128 == 0
    # the same graph alone gives the same count (one launch per graph)
    for k in (0, 17, 95):
        assert gpu.Plan(gs[k]).count()[0] == want[k]
    # launch shapes do not change any count
    for sched in (0, 1):
        b.tune(schedule=sched, group_lanes=16)
        assert np.array_equal(b.count()[1], want)


def test_batch_item_sums_and_per_vertex_paths(gpu, orc):
    """A batch's wedge and tiny kernels add each work item's count to its
    graph (the item records carry the graph index); with 64-bit offsets they
    fall back to per-vertex sums.  Both give the oracle's per-graph counts,
    with every tiny-item rule, launch shape and schedule, eagerly and from a
    captured graph, and the autotuner's choice changes none of them."""
    work = gpu.batch_workload(count=64, seed=23, n_lo=100, n_hi=5000)
    gs = [gpu.Graph.from_coo(s, d, n) for s, d, n, _ in work]
    want = np.array([oracle_total(orc, g) for g in gs], np.uint64)
    for tiny in (-1, 1, 2, 0):
        b = gpu.Batch(gs, tiny_items=tiny)
        for opts in ({}, {"wide_offsets": 1}, {"group_lanes": 8, "schedule": 0}, {"wide_offsets": 0}):
            if opts:
                b.tune(**opts)
            for _ in range(2):  # item sums are cleared by each count
                total, per, _ = b.count()
                np.testing.assert_array_equal(per, want)
                assert total == int(want.sum())
        b.capture()
        for _ in range(3):
            np.testing.assert_array_equal(b.count()[1], want)
    b.autotune(budget_ms=300)
    for _ in range(2):
        np.testing.assert_array_equal(b.count()[1], want)


@pytest.mark.parametrize("tau", [-1, 0, 100000])
def test_batch_mixed_extremes(gpu, orc, tau):
    """Empty graphs, a single triangle, cliques, stars and a dense planted
    graph in one batch; tau = 0 sends every tile down the dense path."""
    parts = [
        (np.zeros(0, np.int32), np.zeros(0, np.int32), 5),
        (np.array([1, 2, 2], np.int32), np.array([0, 0, 1], np.int32), 3),
        graphs.complete(40) + (40,),
        graphs.star(300) + (300,),
        gpu.gen_planted(2048, 128, 0.5, 4.0, seed=3) + (2048,),
        graphs.complete(7) + (7,),
    ]
    gs = [gpu.Graph.from_coo(s, d, n) for s, d, n in parts]
    want = np.array([oracle_total(orc, g) for g in gs], np.uint64)
    total, per, _ = gpu.Batch(gs, tau=tau).count()
    np.testing.assert_array_equal(per, want)
    assert want[2] == 40 * 39 * 38 // 6 and want[3] == 0 and want[1] == 1


def test_batch_graphs_may_be_freed(gpu, orc):
    work = gpu.batch_workload(count=10, seed=3, n_lo=100, n_hi=500)
    gs = [gpu.Graph.from_coo(s, d, n) for s, d, n, _ in work]
    want = [oracle_total(orc, g) for g in gs]
    b = gpu.Batch(gs)
    for g in gs:
        g.close()
    assert b.count()[1].tolist() == want


@pytest.mark.parametrize("flags", [0, 1, 3])
def test_batch_from_coo_matches_graphs(gpu, orc, flags):
    """One GPU canonicalisation of many graphs as separate vertex ranges
    (with empty, loop-only and padded graphs among them) counts every graph
    exactly as it counts alone."""
    work = gpu.batch_workload(40, seed=11, n_lo=64, n_hi=3000)
    graphs = [(s, d, n) for s, d, n, _ in work]
    graphs.insert(5, (np.zeros(0, np.int32), np.zeros(0, np.int32), 17))     # no edges
    graphs.insert(9, (np.arange(6, dtype=np.int32), np.arange(6, dtype=np.int32), 6))  # loops only
    s, d, n = graphs[12]
    graphs[12] = (s, d, n + 500)  # isolated vertices after the used ids
    want = []
    for s, d, n in graphs:
        g = gpu.Graph.from_coo(s, d, n, flags)
        rp, c = g.csr()
        want.append(orc.tc_merge(rp, c, threads=2))
    b = gpu.Batch.from_coo(graphs, flags)
    total, per, _ = b.count()
    assert [int(x) for x in per] == want and total == sum(want)
    assert b.info["graphs"] == len(graphs)
    with pytest.raises(gpu.BtcError):
        gpu.Batch.from_coo([(np.array([0, 3], np.int32), np.array([1, 2], np.int32), 3)])  # id 3 >= n


def test_batch_from_coo_many_graphs(gpu, orc):
    """Hundreds of graphs (more than one block of segment starts)."""
    work = gpu.batch_workload(700, seed=13, n_lo=16, n_hi=200)
    graphs = [(s, d, n) for s, d, n, _ in work]
    b = gpu.Batch.from_coo(graphs)
    total, per, _ = b.count()
    want = []
    for s, d, n in graphs:
        rp, c = gpu.Graph.from_coo(s, d, n).csr()
        want.append(orc.tc_merge(rp, c))
    assert [int(x) for x in per] == want and total == sum(want)
