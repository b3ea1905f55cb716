"""GPU parity: every HIP path against the CPU oracle, bit for bit.

Counts are integers, so the bar is exact equality of the u64 total and of
every int32 per-vertex sum.  Small and medium graphs are compared with the
oracle directly; the full-size workloads are checked through properties
that hold at any size (closed forms, path / tile-width / labelling
invariance, sum of per-vertex counts = total, C(d, 2) bounds).
"""
import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu

_cache = {}


def case(name, btc):
    """(src, dst, n) of a named test graph."""
    if name in _cache:
        return _cache[name]
    if name == "karate":
        s, d, n, _ = graphs.karate()
    elif name == "messy":
        s, d = graphs.messy(3001, 60000, 1)
        n = 3001
    elif name == "rmat12":
        s, d = btc.gen_rmat(12, 16, seed=12)
        n = 1 << 12
    elif name == "rmat15":
        s, d = btc.gen_rmat(15, 16, seed=15)
        n = 1 << 15
    elif name == "er":
        s, d = btc.gen_er(4096, 24.0, seed=3)
        n = 4096
    elif name == "stencil3d":
        s, d = btc.gen_stencil3d(20, 17, 9)
        n = 20 * 17 * 9
    elif name == "stencil2d":
        s, d = btc.gen_stencil2d(61, 33)
        n = 61 * 33
    elif name == "planted":
        s, d = btc.gen_planted(4096, 128, 0.5, 4.0, seed=7)
        n = 4096
    elif name == "planted_dense":  # 256-vertex communities at p = 0.9: rows full of hub-core ids
        s, d = btc.gen_planted(2048, 256, 0.9, 2.0, seed=7)
        n = 2048
    elif name == "complete":
        s, d = graphs.complete(150)
        n = 150
    elif name == "star":
        s, d = graphs.star(3000)
        n = 3000
    else:
        raise KeyError(name)
    _cache[name] = (s, d, n)
    return _cache[name]


CASES = ["karate", "messy", "rmat12", "rmat15", "er", "stencil3d", "stencil2d", "planted", "complete", "star"]


def oracle_counts(orc, g):
    rp, c = g.csr()
    t, pr = orc.tc_merge(rp, c, threads=4, per_row=True)
    return t, pr.astype(np.int64)


@pytest.mark.parametrize("K", [32, 64, 128])
@pytest.mark.parametrize("name", ["karate", "messy", "rmat12", "stencil3d", "planted", "star"])
def test_bitbsr_build_matches_oracle(gpu, orc, K, name):
    s, d, n = case(name, gpu)
    g = gpu.Graph.from_coo(s, d, n)
    rp, c = g.csr()
    b = gpu.BitBSR(g, K)
    got = b.to_host()
    want = orc.bitbsr_build(rp, c, K)
    for x, y in zip(got, want):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("K", [32, 64, 128])
@pytest.mark.parametrize("name", CASES)
def test_bitbsr_count_matches_oracle(gpu, orc, K, name):
    s, d, n = case(name, gpu)
    g = gpu.Graph.from_coo(s, d, n)
    want, want_pv = oracle_counts(orc, g)
    b = gpu.BitBSR(g, K)
    for variant in (gpu.VARIANT_BINSEARCH, gpu.VARIANT_DIRECT, gpu.VARIANT_AUTO):
        for tau in (-1, 0, 16 * K):  # default, all tiles dense, all sparse
            for mfma in (False, True):  # dense tiles by AND + popcount / by i8 MFMA
                t, pv, _ = b.count(variant=variant, tau=tau, per_vertex=True, mfma=mfma)
                assert t == want, (variant, tau, mfma)
                np.testing.assert_array_equal(pv, want_pv)
    t, pv, ms = b.count()
    assert t == want and pv is None and ms > 0


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("K,tau", [(64, -1), (32, 0), (128, 0), (64, 0), (128, 100000)])
def test_engine_matches_oracle(gpu, orc, name, K, tau):
    s, d, n = case(name, gpu)
    g = gpu.Graph.from_coo(s, d, n)
    want, want_pv = oracle_counts(orc, g)
    p = gpu.Plan(g, tile_k=K, tau=tau)
    t, pv, _ = p.count(per_vertex=True)
    assert t == want
    np.testing.assert_array_equal(pv, want_pv)
    t2, pv2, _ = p.count()  # a plan is reusable; totals-only launch
    assert t2 == want and pv2 is None
    for form in (1, 2):  # dense tiles by AND + popcount, then by i8 MFMA
        p.tune(dense_kernel=form)
        t3, pv3, _ = p.count(per_vertex=True)
        assert t3 == want, form
        np.testing.assert_array_equal(pv3, want_pv)
    info = p.info
    assert info["sparse_edges"] + info["dense_edges"] == g.m


@pytest.mark.parametrize("lookup", [0, 1, 2])
@pytest.mark.parametrize("name", ["rmat15", "messy", "planted", "stencil3d"])
def test_engine_lookup_modes(gpu, orc, name, lookup):
    """Bitmap window, LDS hash and HBM binary search give the same counts."""
    s, d, n = case(name, gpu)
    for flags in (gpu.CANON_COMPACT, gpu.CANON_DEFAULT):
        g = gpu.Graph.from_coo(s, d, n, flags)
        want, want_pv = oracle_counts(orc, g)
        t, pv, _ = gpu.Plan(g, lookup=lookup, tau=100000).count(per_vertex=True)
        assert t == want
        np.testing.assert_array_equal(pv, want_pv)


def test_engine_launch_shapes(gpu, orc):
    """Every group width / unroll / schedule / occupancy cap counts the same."""
    for name in ("rmat15", "messy", "stencil2d"):
        s, d, n = case(name, gpu)
        g = gpu.Graph.from_coo(s, d, n)
        want, want_pv = oracle_counts(orc, g)
        p = gpu.Plan(g)
        for G in (4, 8, 16, 32):
            for U in (1, 2, 4):
                for sched in (0, 1):
                    p.tune(group_lanes=G, unroll=U, schedule=sched, waves_per_eu=(0 if U != 2 else 1))
                    t, pv, _ = p.count(per_vertex=(G in (4, 16)))
                    assert t == want, (name, G, U, sched)
                    if pv is not None:
                        np.testing.assert_array_equal(pv, want_pv)


def test_plan_tune_rejects_bad_options(gpu):
    g = gpu.Graph.from_coo(np.array([1, 2, 2], np.int32), np.array([0, 0, 1], np.int32))
    p = gpu.Plan(g)
    for bad in ({"group_lanes": 12}, {"unroll": 3}, {"schedule": 2}, {"lookup": 5}):
        with pytest.raises(gpu.BtcError):
            p.tune(**bad)
    assert p.count()[0] == 1


@pytest.mark.parametrize("name", ["rmat15", "rmat12", "planted_dense", "complete"])
def test_engine_core_bitmap_path(gpu, orc, name):
    """Lists whose rows hold many hub-core ids (< the core size) are counted by a
    bitmap AND + popcount for the core part; the counts match the oracle and
    the path is really taken."""
    s, d, n = case(name, gpu)
    g = gpu.Graph.from_coo(s, d, n)
    want, want_pv = oracle_counts(orc, g)
    p = gpu.Plan(g, tau=100000)
    assert p.info["core_lists"] > 0 and p.info["core_rows"] > 0
    t, pv, _ = p.count(per_vertex=True)
    assert t == want
    np.testing.assert_array_equal(pv, want_pv)
    for G in (8, 16, 32):
        p.tune(group_lanes=G, unroll=2)
        assert p.count()[0] == want


@pytest.mark.parametrize("planner", [1, 2])
def test_engine_16_and_32_bit_prefix_reads(gpu, orc, planner):
    """Pivots whose ids are all below 2^16 read prefixes from the 16-bit copy
    of L's columns, the others from the 32-bit ids: on a graph with
    communities spread over 2^17 ids both kinds count, and the totals and
    per-vertex counts match the oracle."""
    s, d = gpu.gen_planted(1 << 17, 128, 0.5, 4.0, seed=11)
    g = gpu.Graph.from_coo(s, d, 1 << 17)
    want, want_pv = oracle_counts(orc, g)
    rowptr, col = g.csr()
    last = col[rowptr[1:][np.diff(rowptr) > 0] - 1]  # each pivot's largest id
    assert (last < 1 << 16).any() and (last >= 1 << 16).any()
    for tau in (-1, 100000):
        p = gpu.Plan(g, tau=tau, planner=planner)
        t, pv, _ = p.count(per_vertex=True)
        assert t == want
        np.testing.assert_array_equal(pv, want_pv)


@pytest.mark.parametrize("K", [32, 64])
def test_engine_dense_matching_without_column_set(gpu, orc, K):
    """Every tile dense (tau = 0) on a graph with 1.5 M vertices: at K = 32 it
    has more column blocks than the dense kernel's LDS column set holds, so
    tile columns are matched by search; at K = 64 through the column set.
    Both forms, totals and per-vertex counts match the oracle."""
    n = 1_500_000
    s, d = gpu.gen_planted(n, 16, 0.5, 1.0, seed=5)  # 93,750 communities of 16
    g = gpu.Graph.from_coo(s, d, n)
    ncb = (g.n + K - 1) // K  # column blocks
    assert ncb > 43008 if K == 32 else ncb <= 43008
    want, want_pv = oracle_counts(orc, g)
    p = gpu.Plan(g, tile_k=K, tau=0)
    assert p.info["dense_tiles"] >= 16384  # waves take runs of tiles
    for form in (1, 2):
        p.tune(dense_kernel=form)
        t, pv, _ = p.count(per_vertex=True)
        assert t == want, form
        np.testing.assert_array_equal(pv, want_pv)


@pytest.mark.parametrize("planner", [1, 2])
def test_engine_tiny_pivot_items(gpu, orc, planner):
    """Pivots with few lists, at most 64 ids and ids spread over more than a
    wave's bitmap window are counted by groups of 8 lanes in their own
    kernel (beside the wedge kernel for large plans); totals, per-vertex
    counts, shares, every lookup mode and replays match the oracle."""
    s, d = gpu.gen_planted(1 << 17, 128, 0.5, 4.0, seed=11)
    g = gpu.Graph.from_coo(s, d, 1 << 17)
    want, want_pv = oracle_counts(orc, g)
    # by default only graphs with 2^30 wedges or more get tiny items
    assert gpu.Plan(g, planner=planner).info["tiny_items"] == 0
    assert gpu.Plan(g, planner=planner, tiny_items=-1).info["tiny_items"] == 0
    p = gpu.Plan(g, planner=planner, tiny_items=1)
    info = p.info
    assert 0 < info["tiny_items"] < info["work_items"]
    t, pv, _ = p.count(per_vertex=True)
    assert t == want
    np.testing.assert_array_equal(pv, want_pv)
    for k in (2, 3):
        assert sum(p.count_share(k, j)[0] for j in range(k)) == want, k
    for lookup in (1, 2):
        p.tune(lookup=lookup)
        assert p.count()[0] == want, lookup
    p.tune(lookup=0)
    p.capture()
    for _ in range(3):
        assert p.count()[0] == want
        p.count(per_vertex=True)
    # no dense tiles; and a smaller graph (plans below 2^24 checks run their
    # kernels one after another on one stream)
    q = gpu.Plan(g, planner=planner, tau=100000, tiny_items=1)
    assert q.count()[0] == want
    s2, d2 = gpu.gen_rmat(16, 8, seed=16)
    g2 = gpu.Graph.from_coo(s2, d2, 1 << 16)
    want2, want2_pv = oracle_counts(orc, g2)
    t2, pv2, _ = gpu.Plan(g2, planner=planner, tiny_items=1).count(per_vertex=True)
    assert t2 == want2
    np.testing.assert_array_equal(pv2, want2_pv)


@pytest.mark.parametrize("planner", [1, 2])
def test_engine_tiny_items_all_small_pivots(gpu, orc, planner):
    """tiny_items 2 (the batches' default) gives every pivot with few lists
    and at most 64 ids to the tiny kernel, whatever the spread of its ids:
    more tiny items than tiny_items 1, and the same totals and per-vertex
    counts as the oracle, alone, in shares and replayed."""
    for s, d, n in (gpu.gen_planted(1 << 17, 128, 0.5, 4.0, seed=11) + (1 << 17,),
                    gpu.gen_rmat(16, 8, seed=16) + (1 << 16,)):
        g = gpu.Graph.from_coo(s, d, n)
        want, want_pv = oracle_counts(orc, g)
        p1 = gpu.Plan(g, planner=planner, tiny_items=1)
        p2 = gpu.Plan(g, planner=planner, tiny_items=2)
        assert p2.info["tiny_items"] > p1.info["tiny_items"]
        assert p2.info["work_items"] == p1.info["work_items"]
        t, pv, _ = p2.count(per_vertex=True)
        assert t == want
        np.testing.assert_array_equal(pv, want_pv)
        assert sum(p2.count_share(3, j)[0] for j in range(3)) == want
        p2.capture()
        for _ in range(2):
            assert p2.count()[0] == want


@pytest.mark.parametrize("planner", [1, 2])
@pytest.mark.parametrize("core_ids", [-1, 128, 1024, 8192, 32768])
def test_engine_core_size(gpu, orc, core_ids, planner):
    """Any hub-core size (none, one 128-id block, up to a wave's window span)
    gives the same counts, with both planners."""
    s, d, n = case("rmat15", gpu)
    g = gpu.Graph.from_coo(s, d, n)
    want, want_pv = oracle_counts(orc, g)
    p = gpu.Plan(g, tau=100000, core_ids=core_ids, planner=planner)
    info = p.info
    assert info["core_ids"] == max(core_ids, 0)
    assert (info["core_lists"] > 0) == (core_ids > 0)
    t, pv, _ = p.count(per_vertex=True)
    assert t == want
    np.testing.assert_array_equal(pv, want_pv)
    for lookup in (1, 2):
        p.tune(lookup=lookup)
        assert p.count()[0] == want


def test_engine_core_size_rejects(gpu):
    g = gpu.Graph.from_coo(*case("karate", gpu)[:2])
    for bad in (100, 65536, 32768 + 128):
        with pytest.raises(gpu.BtcError):
            gpu.Plan(g, core_ids=bad)
    for bad in (-2, 3):
        with pytest.raises(gpu.BtcError):
            gpu.Plan(g, tiny_items=bad)


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
def test_engine_any_labelling(gpu, orc, flags):
    s, d, n = case("rmat15", gpu)
    g = gpu.Graph.from_coo(s, d, n, flags)
    want, want_pv = oracle_counts(orc, g)
    t, pv, _ = gpu.Plan(g).count(per_vertex=True)
    assert t == want
    np.testing.assert_array_equal(pv, want_pv)


@pytest.mark.parametrize("chunk", [1, 64, 4096])
def test_engine_work_item_size_invariant(gpu, orc, chunk):
    s, d, n = case("rmat15", gpu)
    g = gpu.Graph.from_coo(s, d, n)
    want, want_pv = oracle_counts(orc, g)
    t, pv, _ = gpu.Plan(g, chunk_work=chunk).count(per_vertex=True)
    assert t == want
    np.testing.assert_array_equal(pv, want_pv)


def test_engine_long_pivot_rows(gpu, orc):
    """Pivot rows longer than the LDS hash table (2048) take the HBM binary
    search: vertex 3000 sees 0..2999, vertices 3001.. see 3000 and a part of
    0..2999; id order keeps d_L(3000) = 3000."""
    rng = np.random.default_rng(5)
    src = [np.full(3000, 3000), ]
    dst = [np.arange(3000)]
    for v in range(3001, 3040):
        nb = rng.choice(3000, 700, replace=False)
        src += [np.full(701, v)]
        dst += [np.concatenate([[3000], nb])]
    s = np.concatenate(src).astype(np.int32)
    d = np.concatenate(dst).astype(np.int32)
    g = gpu.Graph.from_coo(s, d, 3040, gpu.CANON_COMPACT)
    assert g.info["max_row"] >= 3000
    want, want_pv = oracle_counts(orc, g)
    t, pv, _ = gpu.Plan(g, tau=100000).count(per_vertex=True)
    assert t == want
    np.testing.assert_array_equal(pv, want_pv)


def test_empty_and_tiny_graphs(gpu):
    g = gpu.Graph.from_coo(np.zeros(0, np.int32), np.zeros(0, np.int32), n=5)
    assert gpu.Plan(g).count()[0] == 0
    assert gpu.BitBSR(g, 32).count()[0] == 0
    g = gpu.Graph.from_coo(np.array([1], np.int32), np.array([0], np.int32))
    assert gpu.Plan(g).count(per_vertex=True)[0] == 0
    g = gpu.Graph.from_coo(np.array([1, 2, 2], np.int32), np.array([0, 0, 1], np.int32))
    t, pv, _ = gpu.Plan(g).count(per_vertex=True)
    assert t == 1 and pv.tolist() == [0, 0, 1]
    for K in (32, 64, 128):
        assert gpu.BitBSR(g, K).count()[0] == 1


# ---------------------------------------------------------------- full sizes

def test_er_16384_full(gpu, orc):
    """Erdos-Renyi n = 16384, average degree 16: engine and the Lite-policy
    Bit-BSR path equal the oracle exactly."""
    s, d = gpu.gen_er(16384, 16.0, seed=1)
    g = gpu.Graph.from_coo(s, d, 16384)
    want, want_pv = oracle_counts(orc, g)
    K = gpu.lite_tile_k(g.n, g.m)
    assert K == 32
    t, pv, _ = gpu.BitBSR(g, K).count(per_vertex=True)
    assert t == want
    np.testing.assert_array_equal(pv, want_pv)
    t, pv, _ = gpu.Plan(g).count(per_vertex=True)
    assert t == want
    np.testing.assert_array_equal(pv, want_pv)


def test_stencil27_64cubed_full(gpu):
    """27-point stencil on 64^3 in natural order: T = 11,049,696 (closed
    form) on both tile widths of the reference path and on the engine."""
    s, d = gpu.gen_stencil3d(64, 64, 64)
    g = gpu.Graph.from_coo(s, d, 64 ** 3, gpu.CANON_COMPACT)
    want = graphs.stencil27_triangles(64, 64, 64)
    assert want == 11049696
    for K in (32, 128):
        assert gpu.BitBSR(g, K).count()[0] == want
    t, pv, _ = gpu.Plan(g).count(per_vertex=True)
    assert t == want and int(pv.astype(np.int64).sum()) == want


def test_rmat_scale20_properties(gpu):
    """R-MAT scale 20, edge factor 16: engine totals agree across labellings
    and with the 16x128 reference path; per-vertex sums add up to the total
    and respect C(d_L(i), 2)."""
    s, d = gpu.gen_rmat(20, 16, seed=1)
    g = gpu.Graph.from_coo(s, d, 1 << 20)
    t, pv, _ = gpu.Plan(g).count(per_vertex=True)
    assert int(pv.astype(np.int64).sum()) == t
    rp, _ = g.csr()
    deg = np.diff(rp)
    assert np.all(pv.astype(np.int64) <= deg * (deg - 1) // 2)
    g_id = gpu.Graph.from_coo(s, d, 1 << 20, gpu.CANON_COMPACT)
    assert gpu.Plan(g_id).count()[0] == t
    assert gpu.BitBSR(g, 128).count()[0] == t
    assert gpu.count_cpu(g)[0] == t
