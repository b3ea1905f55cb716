"""The captured build + count (btc_pipeline) against the eager paths and the oracle.

One pipeline per capacity replays its HIP graph for every input.  Each
replay's total must equal the oracle's count of the host canonicaliser's
graph, its canonical graph must equal that graph bit for bit (CSR, labels,
statistics), and the eager plan of the same graph must agree; replays of
inputs of different sizes alternate with eager work on the device.
"""
import numpy as np
import pytest

from test_gpu_device_path import same_graph
from test_gpu_parity import CASES, case, oracle_counts

pytestmark = pytest.mark.gpu

_pipes = {}


def pipe(gpu, flags, n_cap=1 << 15, nnz_cap=600000, **opts):
    key = (flags, n_cap, nnz_cap, tuple(sorted(opts.items())))
    if key not in _pipes:
        _pipes[key] = gpu.Pipeline(n_cap, nnz_cap, flags, **opts)
    return _pipes[key]


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_pipeline_matches_host_and_oracle(gpu, orc, flags):
    p = pipe(gpu, flags)
    assert p.info["graph_nodes"] > 20
    for name in CASES:  # every size replays the one capture
        s, d, n = case(name, gpu)
        print("pipeline", name, flags, flush=True)
        host = gpu.Graph.from_coo(s, d, n, flags)
        want, _ = oracle_counts(orc, host)
        t, ms = p.count(s, d, n)
        assert t == want, name
        assert ms > 0
        same_graph(host, p.graph())
        info = p.info
        assert (info["n"], info["m"], info["wedges"]) == (host.n, host.m, host.info["wedges"])
        assert 0 <= info["work_items"] <= host.m  # a star has no prefix lists at all
        assert gpu.Plan(host).count()[0] == want  # eager work between replays


def test_pipeline_replays_interleaved(gpu, orc):
    """Alternating inputs: a small graph after a large one must not see any
    of the large one's words (counts, flags, labels, bitmaps)."""
    p = pipe(gpu, 3)
    names = ["rmat15", "karate", "planted_dense", "star", "rmat15", "complete", "messy", "karate"]
    want = {}
    for name in set(names):
        s, d, n = case(name, gpu)
        want[name] = oracle_counts(orc, gpu.Graph.from_coo(s, d, n))[0]
    for name in names:
        s, d, n = case(name, gpu)
        assert p.count(s, d, n)[0] == want[name], name


def test_pipeline_messy_and_degenerate_inputs(gpu, orc):
    """Duplicates, both orientations, self-loops, unused ids, n inferred,
    no pairs, only self-loops: the same graph as the host canonicaliser."""
    rng = np.random.default_rng(5)
    s = rng.integers(0, 700, 20000).astype(np.int32) * 3
    d = rng.integers(0, 700, 20000).astype(np.int32) * 3
    s[:500] = d[:500]
    s = np.concatenate([s, d[:3000]])
    d = np.concatenate([d, s[:3000]])
    e = np.zeros(0, np.int32)
    loops = np.arange(7, dtype=np.int32)
    for flags in (0, 3):
        p = pipe(gpu, flags)
        for src, dst, n in ((s, d, 0), (s, d, 5000), (e, e, 10), (e, e, 0), (loops, loops, 0)):
            host = gpu.Graph.from_coo(src, dst, n, flags)
            want, _ = oracle_counts(orc, host)
            assert p.count(src, dst, n)[0] == want
            same_graph(host, p.graph())


def test_pipeline_refuses_bad_input(gpu, orc):
    p = pipe(gpu, 3, n_cap=1000, nnz_cap=5000)
    ok_s, ok_d = np.array([0, 1, 2], np.int32), np.array([1, 2, 0], np.int32)
    with pytest.raises(gpu.BtcError):
        p.count(np.array([0, -1], np.int32), np.array([1, 2], np.int32), 5)
    with pytest.raises(gpu.BtcError):
        p.count(np.array([0, 9], np.int32), np.array([1, 2], np.int32), 5)
    with pytest.raises(gpu.BtcError):
        p.count(np.array([0, 1000], np.int32), np.array([1, 2], np.int32), 0)  # n inferred > n_cap
    with pytest.raises(gpu.BtcError):
        p.count(ok_s, ok_d, 1001)  # n > n_cap
    with pytest.raises(gpu.BtcError):
        p.count(np.zeros(5001, np.int32), np.ones(5001, np.int32), 10)  # nnz > nnz_cap
    with pytest.raises(gpu.BtcError):
        p.graph()  # no accepted run since the refusals
    assert p.count(ok_s, ok_d, 3)[0] == 1  # a refusal leaves the pipeline usable
    assert p.count(ok_s, ok_d, 0)[0] == 1


def test_pipeline_device_pairs_and_launch_shapes(gpu, orc):
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    s, d = gpu.gen_rmat(14, 16, seed=3)
    host = gpu.Graph.from_coo(s, d, 1 << 14)
    want, _ = oracle_counts(orc, host)
    ps, pd = C.c_void_p(), C.c_void_p()
    nb = C.c_size_t(s.nbytes)
    assert hip.hipMalloc(C.byref(ps), nb) == 0 and hip.hipMalloc(C.byref(pd), nb) == 0
    try:
        assert hip.hipMemcpy(ps, s.ctypes.data_as(C.c_void_p), nb, 1) == 0
        assert hip.hipMemcpy(pd, d.ctypes.data_as(C.c_void_p), nb, 1) == 0
        for opts in ({}, {"core_ids": -1}, {"lookup": 1}, {"lookup": 2}, {"schedule": 0},
                     {"group_lanes": 16, "unroll": 2}, {"wide_offsets": 1}):
            print("pipeline opts", opts, flush=True)
            p = gpu.Pipeline(1 << 14, s.size, 3, **opts)
            assert p.count_device(ps.value, pd.value, s.size, 1 << 14)[0] == want, opts
            assert p.count(s, d, 1 << 14)[0] == want, opts
            p.close()
    finally:
        hip.hipFree(ps)
        hip.hipFree(pd)


def test_pipeline_rmat18(gpu, orc):
    """R-MAT s18 (4.2M pairs): total and canonical graph as the host path."""
    s, d = gpu.gen_rmat(18, 16, seed=18)
    host = gpu.Graph.from_coo(s, d, 1 << 18)
    want, _ = oracle_counts(orc, host)
    p = gpu.Pipeline(1 << 18, s.size, 3)
    for _ in range(2):
        assert p.count(s, d, 1 << 18)[0] == want
    same_graph(host, p.graph())
    p.close()


def test_pipeline_export_is_independent(gpu, orc):
    """An exported graph and plan count like any plan (totals, per-vertex
    sums, replay, multi-plan launches) and keep doing so after the pipeline
    runs other inputs."""
    p = pipe(gpu, 3)
    kept = []
    for name in ["rmat15", "planted_dense", "karate", "star"]:
        s, d, n = case(name, gpu)
        want, want_pv = oracle_counts(orc, gpu.Graph.from_coo(s, d, n))
        assert p.count(s, d, n)[0] == want
        g, plan = p.export()
        kept.append((name, g, plan, want, want_pv))
    for name, g, plan, want, want_pv in kept:  # the pipeline has run other graphs since
        t, pv, _ = plan.count(per_vertex=True)
        assert t == want, name
        np.testing.assert_array_equal(pv, want_pv)
        info = plan.info
        assert info["dense_tiles"] == 0 and info["work_items"] <= g.m
        plan.capture()
        assert plan.count()[0] == want
    b = gpu.Batch([k[1] for k in kept])  # the exported graphs as ordinary graphs
    np.testing.assert_array_equal(b.count()[1], [k[3] for k in kept])


@pytest.mark.parametrize("nnz", [1, 2047, 2048, 2049, 4095, 4096, 4097, 8191, 12289])
def test_pipeline_tile_boundaries(gpu, orc, nnz):
    """Pair counts around the sort's 2048-key tiles and the scan's 4096-id
    tiles, ids around 1024 / 4096, duplicates and loops mixed in: one
    pipeline replays them all and matches the host canonicaliser and the
    oracle every time."""
    p = pipe(gpu, 3, n_cap=1 << 13, nnz_cap=16384)
    rng = np.random.default_rng(nnz)
    for n in (1023, 1024, 1025, 4097, 8192):
        s = rng.integers(0, n, nnz).astype(np.int32)
        d = rng.integers(0, n, nnz).astype(np.int32)
        d[: nnz // 7] = s[: nnz // 7]  # self-loops
        s[nnz // 2:] = np.minimum(s[nnz // 2:], 40)  # a hub and many duplicates
        host = gpu.Graph.from_coo(s, d, n, 3)
        want, _ = oracle_counts(orc, host)
        assert p.count(s, d, n)[0] == want, (nnz, n)
        same_graph(host, p.graph())


def group_graphs(gpu, k, seed):
    """k small graphs of several kinds, an empty one and one of loops only."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(k):
        n = int(rng.integers(64, 3000))
        kind = i % 5
        if kind == 0:
            s, d = gpu.gen_er(n, float(rng.uniform(4, 16)), seed=seed * 100 + i)
        elif kind == 1:
            sc = int(np.ceil(np.log2(n)))
            s, d = gpu.gen_rmat(sc, 12, seed=seed * 100 + i)
            keep = (s < n) & (d < n)
            s, d = s[keep], d[keep]
        elif kind == 2:
            s, d = gpu.gen_planted(n, 48, 0.5, 2.0, seed=seed * 100 + i)
        elif kind == 3:
            s = d = np.zeros(0, np.int32)
        else:
            s = d = np.arange(n, dtype=np.int32)  # loops only
        out.append((np.asarray(s, np.int32), np.asarray(d, np.int32), n if i % 2 == 0 else 0))
    return out


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_pipeline_batch_matches_oracle(gpu, orc, flags):
    """Several graphs per run: each graph's count equals the oracle's and
    btc_batch_from_coo's, over runs of different group sizes; the exported
    batch counts the same and is untouched by later runs."""
    p = gpu.Pipeline(1 << 16, 1 << 20, flags, max_graphs=16)
    pool = group_graphs(gpu, 24, seed=flags + 1)
    want = [oracle_counts(orc, gpu.Graph.from_coo(s, d, n))[0] for s, d, n in pool]
    rng = np.random.default_rng(9)
    kept = None
    for k in (16, 1, 7, 16, 3):
        pick = rng.choice(len(pool), size=k, replace=False)
        gs = [pool[i] for i in pick]
        t, per, _ = p.count_batch(gs)
        np.testing.assert_array_equal(per, [want[i] for i in pick])
        assert t == sum(want[i] for i in pick)
        if kept is None and k == 7:
            kept = (p.export_batch(), [want[i] for i in pick])
            np.testing.assert_array_equal(kept[0].count()[1], kept[1])
    b = gpu.Batch.from_coo([pool[i] for i in range(8)], flags=flags)
    np.testing.assert_array_equal(b.count()[1], want[:8])
    np.testing.assert_array_equal(kept[0].count()[1], kept[1])  # after later runs
    # single counts on a several-graph pipeline run as one range
    s, d, n = pool[1]
    assert p.count(s, d, n)[0] == want[1]


@pytest.mark.parametrize("opts", [{"wide_offsets": 1}, {"group_lanes": 8, "schedule": 0}])
def test_pipeline_batch_sum_paths(gpu, orc, opts):
    """Per-graph counts of a several-graph run through the per-vertex
    fallback (64-bit offsets) and through item sums with another launch
    shape equal the oracle's, replay after replay and after export."""
    p = gpu.Pipeline(1 << 16, 1 << 20, 3, max_graphs=8, **opts)
    pool = group_graphs(gpu, 10, seed=31)
    want = [oracle_counts(orc, gpu.Graph.from_coo(s, d, n))[0] for s, d, n in pool]
    for pick in ((0, 1, 2, 3, 4, 5, 6, 7), (9, 8), (3,), (2, 4, 6, 8, 9)):
        t, per, _ = p.count_batch([pool[i] for i in pick])
        np.testing.assert_array_equal(per, [want[i] for i in pick])
        assert t == sum(want[i] for i in pick)
    b = p.export_batch()
    np.testing.assert_array_equal(b.count()[1], [want[i] for i in (2, 4, 6, 8, 9)])


def test_pipeline_batch_refuses_bad_input(gpu):
    p = gpu.Pipeline(4096, 1 << 16, 3, max_graphs=4)
    ok = (np.array([0, 1, 2], np.int32), np.array([1, 2, 0], np.int32), 3)
    with pytest.raises(gpu.BtcError):
        p.count_batch([ok, (np.array([0, 5], np.int32), np.array([1, 2], np.int32), 4)])  # id 5 >= 4
    with pytest.raises(gpu.BtcError):
        p.count_batch([ok] * 5)  # more graphs than the pipeline takes
    with pytest.raises(gpu.BtcError):
        p.count_batch([ok, (ok[0], ok[1], 4094)])  # 4097 ids in all
    with pytest.raises(gpu.BtcError):
        gpu.Pipeline(1 << 27, 1 << 16, 3, max_graphs=64)  # 27 + 6 key bits
    t, per, _ = p.count_batch([ok, ok])
    assert t == 2 and list(per) == [1, 1]
