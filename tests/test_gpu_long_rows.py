"""Graphs whose longest L-row exceeds 65536, in the reference's id order.

A hub with the largest id is joined to every other vertex, so its row of L
holds them all (66000 > 65536 ids); random edges among the others close a
triangle with the hub each.  Per-vertex counts then need 64-bit slots on the
device (a vertex's count is at most C(d, 2), which leaves int32 above d =
65536).  Totals and per-vertex counts of the engine, of the
reference-semantics kernels at every tile width and of a batch holding the
graph equal the oracle's; int32 and int64 per-vertex output agree.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_OTHERS = 66000


def hub_graph(seed=1, others=N_OTHERS, degree=6):
    rng = np.random.default_rng(seed)
    m = others * degree // 2
    u = rng.integers(0, others, m).astype(np.int32)
    v = rng.integers(0, others, m).astype(np.int32)
    hub = np.full(others, others, np.int32)
    spokes = np.arange(others, dtype=np.int32)
    return np.concatenate([u, hub]), np.concatenate([v, spokes]), others + 1


@pytest.fixture(scope="module")
def hub(orc):
    s, d, n = hub_graph()
    rp, col, _ = orc.canon(s, d, n, 1)  # compaction only: id order
    want, rows = orc.tc_merge(rp, col, threads=16, per_row=True)
    assert int(np.diff(rp).max()) == N_OTHERS
    return s, d, n, rp, col, int(want), rows


def test_engine_long_row(gpu, hub):
    s, d, n, rp, col, want, rows = hub
    g = gpu.Graph.from_coo(s, d, n, gpu.CANON_COMPACT)
    assert g.info["max_row"] == N_OTHERS
    rp_g, col_g = g.csr()
    assert np.array_equal(rp_g, rp) and np.array_equal(col_g, col)
    for K in (32, 64, 128):
        p = gpu.Plan(g, tile_k=K)
        assert p.count()[0] == want
        t32, pv32, _ = p.count(per_vertex=True)
        t64, pv64, _ = p.count(per_vertex=True, wide=True)
        assert t32 == t64 == want
        np.testing.assert_array_equal(pv32, rows)
        np.testing.assert_array_equal(pv64, rows)
    # device canonicalisation and the device planner
    gd = gpu.Graph.from_coo(s, d, n, gpu.CANON_COMPACT | gpu.CANON_DEVICE)
    t, pv, _ = gpu.Plan(gd, planner=2).count(per_vertex=True, wide=True)
    assert t == want
    np.testing.assert_array_equal(pv, rows)


@pytest.mark.parametrize("K", [32, 64, 128])
def test_plain_long_row(gpu, hub, K):
    s, d, n, rp, col, want, rows = hub
    g = gpu.Graph.from_csr(rp, col)
    b = gpu.BitBSR(g, K)
    for mfma in (False, True):
        t, pv, _ = b.count(per_vertex=True, mfma=mfma)
        assert t == want
        np.testing.assert_array_equal(pv, rows)
        t, pv, _ = b.count(per_vertex=True, mfma=mfma, wide=True, tau=0)  # every tile dense
        assert t == want
        np.testing.assert_array_equal(pv, rows)


def test_batch_with_long_row(gpu, orc, hub):
    s, d, n, rp, col, want, rows = hub
    small = gpu.batch_workload(6, seed=4, n_lo=100, n_hi=900)
    graphs = [(a, b, m) for a, b, m, _ in small]
    graphs.insert(3, (s, d, n))
    expect = []
    for a, b, m in graphs:
        r, c, _ = orc.canon(a, b, m, 1)
        expect.append(orc.tc_merge(r, c, threads=16))
    gs = [gpu.Graph.from_coo(a, b, m, gpu.CANON_COMPACT) for a, b, m in graphs]
    total, per, _ = gpu.Batch(gs).count()
    assert [int(x) for x in per] == expect and total == sum(expect)
    total, per, _ = gpu.Batch.from_coo(graphs, gpu.CANON_COMPACT).count()
    assert [int(x) for x in per] == expect
