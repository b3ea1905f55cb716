"""Serving: every answer equals the oracle's count of its graph, whatever it
was batched with; repeats hit the resident set; a tight HBM
budget evicts idle graphs without changing answers; bad input fails only
its own requests."""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def drain(sv, n, timeout=60.0):
    out = []
    t0 = time.time()
    while len(out) < n and time.time() - t0 < timeout:
        out += sv.poll(wait_ms=50)
    assert len(out) == n, "server did not finish in time"
    return out


def make_graphs(gpu, k, seed):
    rng = np.random.default_rng(seed)
    gs = []
    for i in range(k):
        n = int(rng.integers(200, 6000))
        if i % 3 == 0:
            s, d = gpu.gen_er(n, float(rng.uniform(6, 20)), seed=i + 1)
        elif i % 3 == 1:
            sc = int(np.ceil(np.log2(n)))
            s, d = gpu.gen_rmat(sc, 12, seed=i + 1)
            keep = (s < n) & (d < n)
            s, d = s[keep], d[keep]
        else:
            s, d = gpu.gen_planted(n, 96, 0.4, 3.0, seed=i + 1)
        gs.append((s, d, n))
    return gs


def alone(orc, s, d, n):
    """The graph's count by the oracle (its own canonicalisation and merge)."""
    rp, col, _ = orc.canon(s, d, n, 3)
    return orc.tc_merge(rp, col, threads=4)


@pytest.mark.parametrize("canon_flags", [3, 3 | 4])  # host / GPU canonicalisation and planning
def test_server_mix_equals_alone(gpu, orc, canon_flags):
    gs = make_graphs(gpu, 40, seed=1)
    want = [alone(orc, *g) for g in gs]
    sv = gpu.Server(max_batch=16, prep_threads=4, canon_flags=canon_flags)
    ids = {}
    for k, (s, d, n) in enumerate(gs):
        ids[sv.submit(k, s, d, n)] = k
    # repeats of resident (or in-preparation) graphs, no edges sent
    for k in list(range(0, 40, 3)):
        ids[sv.submit(k)] = k
    res = drain(sv, len(ids))
    for r in res:
        assert r["status"] == 0
        assert r["triangles"] == want[ids[r["request_id"]]]
        assert r["t_submit"] <= r["t_start"] <= r["t_done"]
    st = sv.stats()
    assert st["completed"] == len(ids) and st["prepared"] == 40 and st["resident_hits"] == 14
    assert st["launches"] >= 1 and st["failed"] == 0
    sv.close()


def test_server_eviction_keeps_answers(gpu, orc):
    gs = make_graphs(gpu, 12, seed=2)
    want = [alone(orc, *g) for g in gs]
    # host preparation: one graph per resident unit
    sv = gpu.Server(hbm_budget_bytes=3 << 20, max_batch=4, prep_threads=2, canon_flags=3)
    for rnd in range(2):
        ids = {sv.submit(k, *g): k for k, g in enumerate(gs)}
        for r in drain(sv, len(ids)):
            assert r["status"] == 0 and r["triangles"] == want[ids[r["request_id"]]]
    st = sv.stats()
    assert st["evictions"] > 0 and st["resident_bytes"] <= (3 << 20) + (1 << 22)
    sv.close()


def test_server_grouped_units_evict_and_stay_exact(gpu, orc):
    """GPU preparation groups queued graphs into one unit; a tight budget
    evicts whole idle units, and every answer stays exact."""
    gs = make_graphs(gpu, 24, seed=5)
    want = [alone(orc, *g) for g in gs]
    sv = gpu.Server(hbm_budget_bytes=4 << 20, max_batch=64, prep_threads=2)
    for rnd in range(3):
        base = 1000 * rnd  # fresh keys every round: earlier units must make room
        ids = {sv.submit(base + k, *g): k for k, g in enumerate(gs)}
        ids.update({sv.submit(base + k): k for k in range(0, 24, 5)})  # repeats
        for r in drain(sv, len(ids)):
            assert r["status"] == 0 and r["triangles"] == want[ids[r["request_id"]]]
    st = sv.stats()
    assert st["evictions"] > 0 and st["prepared"] == 72 and st["failed"] == 0
    sv.close()


def test_server_bad_graph_fails_alone(gpu, orc):
    sv = gpu.Server(prep_threads=2)
    good = make_graphs(gpu, 3, seed=3)
    want = [alone(orc, *g) for g in good]
    bad = sv.submit(99, np.array([0, 7], np.int32), np.array([1, 2], np.int32), 5)  # id 7 >= n
    ids = {sv.submit(k, *g): k for k, g in enumerate(good)}
    res = {r["request_id"]: r for r in drain(sv, 4)}
    assert res[bad]["status"] == 1
    for rid, k in ids.items():
        assert res[rid]["status"] == 0 and res[rid]["triangles"] == want[k]
    again = sv.submit(99)  # the failed key keeps failing without new edges
    assert drain(sv, 1)[0]["request_id"] == again
    with pytest.raises(gpu.BtcError):
        sv.submit(12345)  # unknown key without a graph
    sv.close()
