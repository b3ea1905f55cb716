"""The engine's dense-tile kernel on units with long matched lists.

Four row blocks (ids 19200-19263) all join every 64th id below 16000, so
they share about 250 tile columns: a unit's matches span several 64-entry
chunks of the list, wrap the per-wave ring of matched pairs and take many
LDS stage batches.  The 64 ids are also a clique (dense mask tiles at every
tile width, whose slabs past the clique are empty and are skipped), and
random edges elsewhere keep the rest sparse.  Totals and per-vertex
counts equal the oracle's at every tile width, both dense forms, every tau
that keeps the tile dense, and with all tiles forced dense.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def long_list_graph(seed=5):
    rng = np.random.default_rng(seed)
    src, dst = [], []
    hubs = np.arange(0, 16000, 64)
    for r in range(19200, 19264):
        src.append(np.full(hubs.size, r))
        dst.append(hubs)
    block = np.arange(19200, 19264)
    a, b = np.meshgrid(block, block)
    keep = a > b
    src.append(a[keep])
    dst.append(b[keep])
    m = 60000
    src.append(rng.integers(0, 20000, m))
    dst.append(rng.integers(0, 20000, m))
    return np.concatenate(src).astype(np.int32), np.concatenate(dst).astype(np.int32), 20000


@pytest.fixture(scope="module")
def graph(gpu, orc):
    s, d, n = long_list_graph()
    rp, col, _ = orc.canon(s, d, n, 0)  # input ids, uncompacted: the two blocks stay aligned
    want, rows = orc.tc_merge(rp, col, threads=16, per_row=True)
    g = gpu.Graph.from_coo(s, d, n, 0)
    rp_g, col_g = g.csr()
    assert np.array_equal(rp_g, rp) and np.array_equal(col_g, col)
    return g, int(want), rows


@pytest.mark.parametrize("K", [32, 64, 128])
@pytest.mark.parametrize("dense_kernel", [1, 2])
@pytest.mark.parametrize("tau", [-1, 0])
def test_dense_units_long_lists(gpu, graph, K, dense_kernel, tau):
    g, want, rows = graph
    p = gpu.Plan(g, tile_k=K, tau=tau, dense_kernel=dense_kernel)
    assert p.info["dense_tiles"] > 0
    total, pv, _ = p.count(per_vertex=True)
    assert total == want
    np.testing.assert_array_equal(pv, rows)
    assert p.count()[0] == want
    assert sum(p.count_share(3, j)[0] for j in range(3)) == want
