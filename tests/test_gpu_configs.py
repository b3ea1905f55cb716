"""Parity at the full sizes of the benchmark workloads, against the oracle.

* R-MAT scale 22, edge factor 16, (0.57, 0.19, 0.19, 0.05), seed 1: the GPU
  canonicalisation equals the oracle's (CSR and labels), and the engine's
  total and per-vertex counts, the reference-semantics 16 x 128 and 16 x 32
  kernels' totals and per-vertex counts equal the oracle merge
  (triangle_count_cpu.h:60-101) on the same L; the engine also counts the
  id-ordered L (the reference's natural order) exactly.
* The batched workload, 4096 graphs with n log-uniform in [256, 16384]:
  every per-graph count of the batch built on the GPU from the pairs, of the
  batch built from host-canonicalised graphs and of the autotuned batch
  equals the oracle's count of that graph.
* The serving stream: 200 Poisson arrivals at a fixed rate, n log-normal
  (median 2^13, capped at 2^20), 30 % repeats; every answer equals the
  oracle's count of its graph.

The oracle canonicalises each graph itself (orc_canon), so every expected
count is independent of the library.
"""
import os
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)  # the box's CPU share


def oracle_count(orc, s, d, n, per_row=False):
    rp, col, _ = orc.canon(s, d, n, 3)
    return orc.tc_merge(rp, col, threads=1, per_row=per_row)


def oracle_counts(orc, graphs):
    """Oracle count of every (src, dst, n) in graphs, THREADS graphs at a time."""
    with ThreadPoolExecutor(THREADS) as ex:
        return list(ex.map(lambda g: oracle_count(orc, g[0], g[1], g[2]), graphs))


@pytest.fixture(scope="module")
def rmat22(gpu, orc):
    print("R-MAT s22 ef16 seed 1: generating and canonicalising", flush=True)
    s, d = gpu.gen_rmat(22, 16, 0.57, 0.19, 0.19, seed=1)
    rp, col, lab = orc.canon(s, d, 1 << 22, 3)
    want, rows = orc.tc_merge(rp, col, threads=THREADS, per_row=True)
    return s, d, rp, col, lab, int(want), rows


@pytest.mark.timeout(900)
def test_rmat22_engine_matches_oracle(gpu, rmat22):
    s, d, rp, col, lab, want, rows = rmat22
    g = gpu.Graph.from_coo(s, d, 1 << 22, gpu.CANON_DEFAULT | gpu.CANON_DEVICE)
    rp_g, col_g = g.csr()
    assert np.array_equal(rp_g, rp) and np.array_equal(col_g, col), "GPU canonicalisation differs from the oracle"
    assert np.array_equal(g.labels(), lab)
    p = gpu.Plan(g)
    info = p.info
    assert info["sparse_edges"] + info["dense_edges"] == g.m
    total, pv, _ = p.count(per_vertex=True)
    assert total == want
    np.testing.assert_array_equal(pv, rows)
    for _ in range(3):  # the totals-only count the benchmark times, repeated
        assert p.count()[0] == want
    p.capture()  # replays of the captured count, with eager work between them
    for _ in range(3):
        assert p.count()[0] == want
        assert p.count(per_vertex=True)[0] == want


@pytest.mark.timeout(900)
@pytest.mark.parametrize("K", [128, 32])
def test_rmat22_plain_matches_oracle(gpu, rmat22, K):
    s, d, rp, col, lab, want, rows = rmat22
    g = gpu.Graph.from_csr(rp, col)
    b = gpu.BitBSR(g, K)
    total, pv, _ = b.count(per_vertex=True)
    assert total == want
    np.testing.assert_array_equal(pv, rows)
    assert b.count(mfma=True)[0] == want


@pytest.mark.timeout(900)
def test_rmat22_id_order_engine(gpu, rmat22):
    """The reference's natural (compacted input) order: same total."""
    s, d, _, _, _, want, _ = rmat22
    g = gpu.Graph.from_coo(s, d, 1 << 22, gpu.CANON_COMPACT | gpu.CANON_DEVICE)
    print("id order: n=%d m=%d max_row=%d" % (g.n, g.m, g.info["max_row"]), flush=True)
    assert gpu.Plan(g).count()[0] == want


@pytest.fixture(scope="module")
def batch4096(gpu, orc):
    print("batched workload: 4096 graphs, generating and counting on the CPU", flush=True)
    work = gpu.batch_workload(4096, seed=1)
    graphs = [(s, d, n) for s, d, n, _ in work]
    want = np.array(oracle_counts(orc, graphs), np.uint64)
    return graphs, want


@pytest.mark.timeout(900)
def test_batch4096_from_pairs_matches_oracle(gpu, batch4096):
    graphs, want = batch4096
    b = gpu.Batch.from_coo(graphs)
    total, per, _ = b.count()
    np.testing.assert_array_equal(per, want)
    assert total == int(want.sum())
    b.capture()
    for _ in range(2):
        np.testing.assert_array_equal(b.count()[1], want)


@pytest.mark.timeout(900)
def test_batch4096_graphs_and_autotune_match_oracle(gpu, batch4096):
    graphs, want = batch4096
    gs = [gpu.Graph.from_coo(s, d, n) for s, d, n in graphs]
    b = gpu.Batch(gs)
    for g in gs:
        g.close()
    np.testing.assert_array_equal(b.count()[1], want)
    best, _ = b.autotune(budget_ms=2000)
    print("autotuned batch options:", best, flush=True)
    np.testing.assert_array_equal(b.count()[1], want)


@pytest.fixture(scope="module")
def serving_stream(gpu, orc):
    print("serving stream: 200 requests, counting every graph on the CPU", flush=True)
    reqs = gpu.serving_workload(200, seed=3, key0=1000)
    graphs = {k: g for k, g in reqs if g is not None}
    keys = sorted(graphs)
    counts = oracle_counts(orc, [graphs[k] for k in keys])
    return reqs, dict(zip(keys, counts))


@pytest.mark.timeout(900)
# GPU preparation captured (pipelines) or eager; host preparation
@pytest.mark.parametrize("canon_flags,captured", [(3 | 4, 1), (3 | 4, 0), (3, 0)])
def test_serving_stream_matches_oracle(gpu, serving_stream, canon_flags, captured):
    reqs, want = serving_stream
    rng = np.random.default_rng(5)
    gaps = rng.exponential(1.0 / 300.0, len(reqs))  # Poisson arrivals, 300 requests/s
    sv = gpu.Server(canon_flags=canon_flags, captured_prep=captured)
    t0 = time.perf_counter()

    def submit():
        t = 0.0
        for (key, g), gap in zip(reqs, gaps):
            t += gap
            delay = t - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            if g is None:
                sv.submit(key)
            else:
                sv.submit(key, g[0], g[1], g[2])

    th = threading.Thread(target=submit)
    th.start()
    res = []
    while len(res) < len(reqs) and time.perf_counter() - t0 < 300:
        res += sv.poll(wait_ms=50)
    th.join()
    assert len(res) == len(reqs), "server did not answer every request"
    for r in res:
        assert r["status"] == 0, r
        assert r["triangles"] == want[r["graph_key"]], r
    st = sv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["resident_hits"] == sum(1 for _, g in reqs if g is None)
    sv.close()
