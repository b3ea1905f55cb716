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
