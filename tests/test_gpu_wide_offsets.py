"""The wedge kernel's 64-bit byte offsets into L.

* Forced (wide_offsets = 1) on ordinary graphs: totals, per-vertex counts
  and the parts of a split equal the oracle's, in every lookup mode.
* A graph whose L holds more than 2^30 ids, where 32-bit byte offsets of
  4-byte ids end: a strip of 2^29 + 2^20 vertices with edges (i, i + 1) and
  (i, i + 2).  The engine takes the 64-bit offsets by itself; the count is
  the closed form n - 2 (the triangles {i, i + 1, i + 2}), the per-vertex
  counts add up to it and none exceeds 3 (a vertex closes at most the three
  triangles it belongs to).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lookup", [0, 1, 2])
def test_forced_wide_offsets_match_oracle(gpu, orc, lookup):
    cases = [gpu.gen_rmat(16, 16, seed=11) + (1 << 16,), gpu.gen_planted(1 << 14, 128, 0.5, 4.0, seed=2) + (1 << 14,),
             gpu.gen_stencil3d(20, 20, 20) + (8000,)]
    for s, d, n in cases:
        g = gpu.Graph.from_coo(s, d, n)
        rp, col = g.csr()
        want, rows = orc.tc_merge(rp, col, threads=16, per_row=True)
        p = gpu.Plan(g, wide_offsets=1, lookup=lookup, tiny_items=1)
        total, pv, _ = p.count(per_vertex=True)
        assert total == want
        np.testing.assert_array_equal(pv, rows)
        assert sum(p.count_share(3, j)[0] for j in range(3)) == want


@pytest.mark.timeout(600)
def test_graph_beyond_2_30_ids(gpu):
    n = (1 << 29) + (1 << 20)
    i = np.arange(n - 1, dtype=np.int32)
    s = np.concatenate([i, i[:-1]])
    d = np.concatenate([i + 1, i[:-1] + 2])
    del i
    print("strip: n=%d pairs=%d" % (n, s.size), flush=True)
    g = gpu.Graph.from_coo(s, d, n, gpu.CANON_DEFAULT | gpu.CANON_DEVICE)
    del s, d
    assert g.n == n and g.m == 2 * n - 3 and g.m + 16 > (1 << 30)
    p = gpu.Plan(g)
    print("plan: %s" % p.info, flush=True)
    assert p.count()[0] == n - 2
    total, pv, _ = p.count(per_vertex=True)
    assert total == n - 2
    assert int(pv.sum(dtype=np.int64)) == n - 2 and int(pv.max()) <= 3 and int(pv.min()) >= 0
    p.close()
    g.close()
