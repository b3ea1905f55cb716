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


def settled(sv, prepared, timeout=30.0):
    """Server stats once `prepared` graphs are resident: with captured
    preparation a graph's requests are answered before its copy is
    published."""
    t0 = time.time()
    st = sv.stats()
    while st["prepared"] < prepared and time.time() - t0 < timeout:
        time.sleep(0.01)
        st = sv.stats()
    return st


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


# host preparation; GPU preparation, eager (grouped units) or captured (pipelines)
@pytest.mark.parametrize("canon_flags,captured", [(3, 0), (3 | 4, 0), (3 | 4, 1)])
def test_server_mix_equals_alone(gpu, orc, canon_flags, captured):
    gs = make_graphs(gpu, 40, seed=1)
    want = [alone(orc, *g) for g in gs]
    sv = gpu.Server(max_batch=16, prep_threads=4, canon_flags=canon_flags, captured_prep=captured)
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
    st = settled(sv, 40)
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


@pytest.mark.parametrize("captured", [0, 1])
def test_server_grouped_units_evict_and_stay_exact(gpu, orc, captured):
    """Eager GPU preparation groups queued graphs into one unit, captured
    preparation makes one unit per graph; a tight budget evicts idle units,
    and every answer stays exact."""
    gs = make_graphs(gpu, 24, seed=5)
    want = [alone(orc, *g) for g in gs]
    sv = gpu.Server(hbm_budget_bytes=4 << 20, max_batch=64, prep_threads=2, captured_prep=captured)
    for rnd in range(3):
        base = 1000 * rnd  # fresh keys every round: earlier units must make room
        ids = {sv.submit(base + k, *g): k for k, g in enumerate(gs)}
        ids.update({sv.submit(base + k): k for k in range(0, 24, 5)})  # repeats
        for r in drain(sv, len(ids)):
            assert r["status"] == 0 and r["triangles"] == want[ids[r["request_id"]]]
    st = settled(sv, 72)
    assert st["evictions"] > 0 and st["prepared"] == 72 and st["failed"] == 0
    sv.close()


@pytest.mark.parametrize("captured", [0, 1])
def test_server_bad_graph_fails_alone(gpu, orc, captured):
    sv = gpu.Server(prep_threads=2, captured_prep=captured)
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
