"""Pins the CPU oracle before it is trusted as the checker (CPU only).

Each oracle function is checked against something it does not restate: the
LAGraph known answer for karate, closed forms, brute force, the reference's
own CPU verifier compiled from its header, set-based canonicalisation on
tiny inputs, and invariances the reference's own tests use (tau sweep).
"""
import itertools

import numpy as np
import pytest

import graphs


def brute_canon(src, dst, n, flags):
    """Set-based canonicalisation (the definition in convert.h:335-380)."""
    edges = {(max(u, v), min(u, v)) for u, v in zip(src.tolist(), dst.tolist()) if u != v}
    ids = list(range(n))
    if flags & 1 and edges:
        used = sorted({x for e in edges for x in e})
        ids = used
    pos = {v: k for k, v in enumerate(ids)}
    edges = {(pos[a], pos[b]) for a, b in edges}
    labels = list(ids)
    if flags & 2:
        deg = [0] * len(ids)
        for a, b in edges:
            deg[a] += 1
            deg[b] += 1
        order = sorted(range(len(ids)), key=lambda v: (-deg[v], v))
        rank = {v: r for r, v in enumerate(order)}
        edges = {(max(rank[a], rank[b]), min(rank[a], rank[b])) for a, b in edges}
        labels = [ids[v] for v in order]
    nn = len(ids)
    rows = [[] for _ in range(nn)]
    for a, b in sorted(edges):
        rows[a].append(b)
    rowptr = np.zeros(nn + 1, np.int64)
    rowptr[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([c for r in rows for c in r], np.int32)
    return rowptr, col, np.array(labels, np.int32)


def test_karate_known_answer(orc):
    s, d, n, t = graphs.karate()
    rp, c, _ = orc.canon(s, d, n, 3)
    assert len(c) == 78
    assert orc.tc_merge(rp, c) == t == 45
    assert orc.tc_brute(rp, c) == 45


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
@pytest.mark.parametrize("seed", range(6))
def test_canon_matches_set_definition(orc, flags, seed):
    s, d = graphs.messy(40, 120, seed)
    n = 45  # ids 40..44 unused: compaction must drop them
    rp, c, lb = orc.canon(s, d, n, flags)
    rp2, c2, lb2 = brute_canon(s, d, n, flags)
    np.testing.assert_array_equal(rp, rp2)
    np.testing.assert_array_equal(c, c2)
    np.testing.assert_array_equal(lb, lb2)


def test_canon_rejects_bad_ids(orc):
    with pytest.raises(ValueError):
        orc.canon(np.array([0, 5], np.int32), np.array([1, 2], np.int32), 4, 3)
    with pytest.raises(ValueError):
        orc.canon(np.array([-1], np.int32), np.array([1], np.int32), 4, 3)


@pytest.mark.parametrize("n", [3, 4, 7, 30])
def test_complete_graph(orc, n):
    s, d = graphs.complete(n)
    rp, c, _ = orc.canon(s, d, 0, 3)
    t, pr = orc.tc_merge(rp, c, per_row=True)
    assert t == n * (n - 1) * (n - 2) // 6
    # in K_n vertex i (as the largest id) closes C(i, 2) triangles
    np.testing.assert_array_equal(pr, [i * (i - 1) // 2 for i in range(n)])


@pytest.mark.parametrize("maker", [lambda: graphs.star(50), lambda: graphs.random_tree(200, 1),
                                   lambda: graphs.bipartite(30, 40, 2)])
def test_triangle_free(orc, maker):
    s, d = maker()
    rp, c, _ = orc.canon(s, d, 0, 3)
    assert orc.tc_merge(rp, c) == 0
    assert orc.tc_brute(rp, c) == 0


@pytest.mark.parametrize("a,b", [(2, 2), (5, 9), (32, 17)])
def test_stencil9_closed_form(orc, btc, a, b):
    s, d = btc.gen_stencil2d(a, b)
    rp, c, _ = orc.canon(s, d, a * b, 3)
    assert orc.tc_merge(rp, c) == graphs.stencil9_triangles(a, b)


@pytest.mark.parametrize("a,b,c", [(2, 2, 2), (3, 5, 7), (8, 8, 8)])
def test_stencil27_closed_form(orc, btc, a, b, c):
    s, d = btc.gen_stencil3d(a, b, c)
    rp, cl, _ = orc.canon(s, d, a * b * c, 1)
    assert orc.tc_merge(rp, cl) == graphs.stencil27_triangles(a, b, c)


@pytest.mark.parametrize("seed", range(5))
@pytest.mark.parametrize("flags", [1, 3])
def test_merge_equals_brute_with_rows(orc, seed, flags):
    s, d = graphs.messy(300, 3000, seed)
    rp, c, _ = orc.canon(s, d, 300, flags)
    t1, p1 = orc.tc_merge(rp, c, per_row=True)
    t2, p2 = orc.tc_brute(rp, c, per_row=True)
    assert t1 == t2
    np.testing.assert_array_equal(p1, p2)


@pytest.mark.parametrize("seed", range(3))
def test_merge_equals_reference_cpu_verifier(orc, seed):
    if not orc.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    s, d = graphs.messy(2000, 30000, seed)
    rp, c, _ = orc.canon(s, d, 2000, 1)
    assert orc.ref_cpu_tc(rp, c, 2) == orc.tc_merge(rp, c, 2)


def test_permutation_invariance(orc):
    s, d = graphs.messy(500, 8000, 7)
    rp, c, _ = orc.canon(s, d, 500, 1)
    t = orc.tc_merge(rp, c)
    perm = np.random.default_rng(3).permutation(500).astype(np.int32)
    rp2, c2, _ = orc.canon(perm[s], perm[d], 500, 1)
    assert orc.tc_merge(rp2, c2) == t
    rp3, c3, _ = orc.canon(s, d, 500, 3)
    assert orc.tc_merge(rp3, c3) == t


def decode_bitbsr(K, n, indptr, indices, rows, tiles):
    """Set of (row, col) bits held by a Bit-BSR."""
    wpt = K // 4
    out = set()
    for t in range(indices.size):
        I, J = int(rows[t]), int(indices[t])
        for w in range(wpt):
            word = int(tiles[t * wpt + w])
            while word:
                b = (word & -word).bit_length() - 1
                word &= word - 1
                pos = w * 64 + b
                out.add((16 * I + pos // K, K * J + pos % K))
    return out


@pytest.mark.parametrize("K", [32, 64, 128])
@pytest.mark.parametrize("seed", range(2))
def test_bitbsr_build_invariants(orc, K, seed):
    s, d = graphs.messy(700, 9000, seed)
    rp, c, _ = orc.canon(s, d, 700, 3)
    n = rp.size - 1
    indptr, indices, rows, tiles = orc.bitbsr_build(rp, c, K)
    n_rb = (n + 15) // 16
    assert indptr[0] == 0 and indptr[-1] == indices.size and np.all(np.diff(indptr.astype(np.int64)) >= 0)
    for I in range(n_rb):
        seg = indices[indptr[I]:indptr[I + 1]]
        assert np.all(np.diff(seg.astype(np.int64)) > 0)
        assert np.all(rows[indptr[I]:indptr[I + 1]] == I)
    # popcount of all tiles = nnz(L), no empty tile, and decoding gives back L
    pops = np.array([bin(int(x)).count("1") for x in tiles]).reshape(-1, K // 4).sum(1)
    assert pops.sum() == c.size and np.all(pops > 0)
    L = {(i, int(j)) for i in range(n) for j in c[rp[i]:rp[i + 1]]}
    assert decode_bitbsr(K, n, indptr, indices, rows, tiles) == L


def test_bitbsr_reference_bit_layout(orc):
    """Bit (r, c) of a 16x128 tile is bit p % 32 of 32-bit word p / 32,
    p = r * 128 + c (bcsr_16x128.h:66-71, 125-128)."""
    # single edge (row 17, col 3): tile (1, 0), r = 1, c = 3 -> p = 131 -> word 4, bit 3
    rp = np.zeros(19, np.int64)
    rp[18:] = 1
    c = np.array([3], np.int32)
    indptr, indices, rows, tiles = orc.bitbsr_build(rp, c, 128)
    words32 = tiles.view(np.uint32)
    assert list(indptr) == [0, 0, 1] and indices.tolist() == [0] and rows.tolist() == [1]
    assert words32[4] == 1 << 3 and words32.sum() == 1 << 3


@pytest.mark.parametrize("K", [32, 64, 128])
@pytest.mark.parametrize("variant", [1, 2])
def test_tile_count_tau_invariant(orc, K, variant):
    """The reference's tau sweep asserts every tau gives the same count
    (apps/tau_sweep.cu:529-537); tau = -1 is all-dense, 16K all-sparse."""
    s, d = btc_free_planted(seed=5)
    rp, c, _ = orc.canon(s, d, 0, 3)
    n = rp.size - 1
    t_ref, p_ref = orc.tc_merge(rp, c, per_row=True)
    bsr = orc.bitbsr_build(rp, c, K)
    seen_dense = set()
    for tau in (-1, 0, 8, K // 2, 16 * K):
        t, pr, nd = orc.bitbsr_count(K, n, *bsr, tau=tau, variant=variant, per_row=True)
        assert t == t_ref
        np.testing.assert_array_equal(pr, p_ref)
        seen_dense.add(nd)
    assert len(seen_dense) > 1  # both paths were exercised


def btc_free_planted(seed):
    """Small planted-community graph built with numpy (dense 16x16 blocks)."""
    rng = np.random.default_rng(seed)
    src, dst = [], []
    for comm in range(6):
        base = comm * 24
        i, j = np.tril_indices(24, -1)
        keep = rng.random(i.size) < 0.6
        src.append(base + i[keep])
        dst.append(base + j[keep])
    u = rng.integers(0, 144, 300)
    v = rng.integers(0, 144, 300)
    return np.concatenate(src + [u]).astype(np.int32), np.concatenate(dst + [v]).astype(np.int32)


def test_tile_width_invariance(orc):
    s, d = graphs.messy(900, 20000, 11)
    rp, c, _ = orc.canon(s, d, 900, 1)
    n = rp.size - 1
    want = orc.tc_merge(rp, c)
    for K in (32, 64, 128):
        assert orc.bitbsr_count(K, n, *orc.bitbsr_build(rp, c, K))[0] == want


def test_lite_policy_table(orc):
    # apps/btc_tc_lite.cu:45-65
    assert orc.lite_tile_k(16384, 131000, 0) == 32      # n < 50k
    assert orc.lite_tile_k(262144, 3298428, 0) == 32    # avg 12.6 < 15
    assert orc.lite_tile_k(2395967, 64156180, 0) == 128  # avg 26.8, n >= 50k
    assert orc.lite_tile_k(60000, 60000 * 15, 0) == 128
    assert orc.lite_tile_k(60000, 60000 * 15 - 1, 0) == 32
    assert orc.lite_tile_k(49999, 10 ** 8, 0) == 32
    assert orc.lite_tile_k(10, 1, 2) == 128 and orc.lite_tile_k(10 ** 6, 10 ** 9, 1) == 32
    # tc_16x128_adaptive.h / tc_16x32_adaptive.h
    assert orc.variant(128, 2047) == 1 and orc.variant(128, 2048) == 2
    assert orc.variant(32, 39999) == 1 and orc.variant(32, 40000) == 2
