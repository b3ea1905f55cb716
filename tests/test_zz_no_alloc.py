"""btc_plan_count_noalloc, the engine count that allocates nothing, called
through the C ABI: its total and per-vertex counts equal btc_plan_count's
and the oracle's on two calls in a row, the second on a smaller graph (the
static workspace keeps the first call's sums, which must not leak into the
second), before and after the plan's side streams exist; sizes beyond the
static workspace and rows beyond int32 per-vertex sums are refused with
their status codes while totals still count.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BTC_ERR_ARG = 1
BTC_ERR_OVERFLOW = 4
PV_MAX_VERTICES = 1 << 22  # 16 MiB of int32 sums in the static workspace


def noalloc(btc, plan, per_vertex):
    """Direct C-ABI call: (status, total, per-vertex array or None)."""
    total = C.c_uint64(0)
    pv = np.full(plan.graph.n, -7, np.int32) if per_vertex else None
    ptr = pv.ctypes.data_as(C.c_void_p) if pv is not None and pv.size else None
    rc = btc.lib().btc_plan_count_noalloc(plan._h, None, C.byref(total), ptr)
    return rc, total.value, pv


def oracle_rows(orc, g):
    rp, col = g.csr()
    return orc.tc_merge(rp, col, threads=16, per_row=True)


def test_noalloc_equals_plan_count_twice(gpu, orc):
    # a large graph whose count forks the dense-tile and tiny-pivot kernels
    # beside the wedge kernel, then a smaller one with dense tiles
    s, d = gpu.gen_rmat(18, 16, seed=7)
    big = gpu.Graph.from_coo(s, d, 1 << 18)
    s2, d2 = gpu.gen_planted(4096, 128, 0.5, 4.0, seed=3)
    small = gpu.Graph.from_coo(s2, d2, 4096)
    assert small.n < big.n
    want_big, rows_big = oracle_rows(orc, big)
    want_small, rows_small = oracle_rows(orc, small)

    pb = gpu.Plan(big)
    print("big: n=%d m=%d info=%s" % (big.n, big.m, pb.info), flush=True)
    # first call on a fresh plan: no side streams yet, the kernels run in turn
    rc, t, pv = noalloc(gpu, pb, True)
    assert rc == 0, gpu.lib().btc_last_error()
    assert t == want_big
    np.testing.assert_array_equal(pv, rows_big)
    # the allocating entry point (creates the side streams) agrees
    t_ref, pv_ref, _ = pb.count(per_vertex=True)
    assert t_ref == t
    np.testing.assert_array_equal(pv_ref, pv)
    # again, now forking onto the plan's side streams
    rc, t, pv = noalloc(gpu, pb, True)
    assert rc == 0 and t == want_big
    np.testing.assert_array_equal(pv, rows_big)
    rc, t, pv = noalloc(gpu, pb, False)
    assert rc == 0 and t == want_big and pv is None

    # second call: a smaller graph in the same workspace
    ps = gpu.Plan(small)
    print("small: n=%d m=%d dense_tiles=%d" % (small.n, small.m, ps.info["dense_tiles"]), flush=True)
    rc, t, pv = noalloc(gpu, ps, True)
    assert rc == 0 and t == want_small
    np.testing.assert_array_equal(pv, rows_small)
    t_ref, pv_ref, _ = ps.count(per_vertex=True)
    assert t_ref == t
    np.testing.assert_array_equal(pv_ref, pv)
    assert ps.count_noalloc()[0] == want_small

    # the captured count of the big plan keeps its own buffers
    pb.capture()
    assert pb.count()[0] == want_big
    rc, t, pv = noalloc(gpu, pb, True)
    assert rc == 0 and t == want_big
    np.testing.assert_array_equal(pv, rows_big)
    assert pb.count()[0] == want_big


def test_noalloc_limits(gpu, orc):
    # a strip of triangles (i, i+1, i+2) over more vertices than the
    # workspace holds sums for: n - 2 triangles
    n = PV_MAX_VERTICES + 5000
    i = np.arange(n - 1, dtype=np.int32)
    s = np.concatenate([i, i[:-1]])
    d = np.concatenate([i + 1, i[:-1] + 2])
    g = gpu.Graph.from_coo(s, d, n)
    assert g.n == n
    p = gpu.Plan(g)
    rc, _, _ = noalloc(gpu, p, True)
    assert rc == BTC_ERR_ARG, rc
    assert b"static workspace" in gpu.lib().btc_last_error()
    rc, t, _ = noalloc(gpu, p, False)
    assert rc == 0 and t == n - 2

    # a hub row of 66000 ids (id order): per-vertex sums need 64-bit slots
    others = 66000
    rng = np.random.default_rng(1)
    u = rng.integers(0, others, 3 * others).astype(np.int32)
    v = rng.integers(0, others, 3 * others).astype(np.int32)
    hs = np.concatenate([u, np.full(others, others, np.int32)])
    hd = np.concatenate([v, np.arange(others, dtype=np.int32)])
    h = gpu.Graph.from_coo(hs, hd, others + 1, gpu.CANON_COMPACT)
    assert h.info["max_row"] == others
    rp, col = h.csr()
    want = orc.tc_merge(rp, col, threads=16)
    ph = gpu.Plan(h)
    rc, _, _ = noalloc(gpu, ph, True)
    assert rc == BTC_ERR_OVERFLOW, rc
    rc, t, _ = noalloc(gpu, ph, False)
    assert rc == 0 and t == want
