"""GPU canonicalisation and the GPU planner against the host path and the oracle.

The device canonicaliser must give the host canonicaliser's L, labels and
statistics bit for bit; the device planner must give exact counts and the
same order-independent plan statistics as the host planner.
"""
import numpy as np
import pytest

import graphs
from test_gpu_parity import CASES, case, oracle_counts

pytestmark = pytest.mark.gpu

INFO_KEYS = ("n", "n_input", "m", "max_row", "wedges", "merge_work")
PLAN_KEYS = ("sparse_edges", "dense_edges", "dense_tiles", "wedge_checks", "sparse_merge_work", "core_lists",
             "core_rows", "tiny_items")


def same_graph(a, b):
    ia, ib = a.info, b.info
    for k in INFO_KEYS:
        assert ia[k] == ib[k], k
    ra, ca = a.csr()
    rb, cb = b.csr()
    np.testing.assert_array_equal(ra, rb)
    np.testing.assert_array_equal(ca, cb)
    np.testing.assert_array_equal(a.labels(), b.labels())


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
@pytest.mark.parametrize("name", CASES)
def test_device_canon_matches_host(gpu, name, flags):
    s, d, n = case(name, gpu)
    host = gpu.Graph.from_coo(s, d, n, flags)
    dev = gpu.Graph.from_coo(s, d, n, flags | gpu.CANON_DEVICE)
    same_graph(host, dev)


@pytest.mark.parametrize("flags", [0, 3])
def test_device_canon_messy_inputs(gpu, flags):
    """Duplicates, both orientations, self-loops, unused ids, n inferred."""
    rng = np.random.default_rng(5)
    s = rng.integers(0, 700, 20000).astype(np.int32) * 3  # every third id used
    d = rng.integers(0, 700, 20000).astype(np.int32) * 3
    s[:500] = d[:500]  # self-loops
    s = np.concatenate([s, d[:3000]])
    d = np.concatenate([d, s[:3000]])  # reversed duplicates
    for n in (0, 5000):
        same_graph(gpu.Graph.from_coo(s, d, n, flags), gpu.Graph.from_coo(s, d, n, flags | gpu.CANON_DEVICE))


def test_device_canon_degenerate(gpu):
    e = np.zeros(0, np.int32)
    for flags in (0, 3):
        same_graph(gpu.Graph.from_coo(e, e, 10, flags), gpu.Graph.from_coo(e, e, 10, flags | gpu.CANON_DEVICE))
        loops = np.arange(7, dtype=np.int32)
        same_graph(gpu.Graph.from_coo(loops, loops, 0, flags),
                   gpu.Graph.from_coo(loops, loops, 0, flags | gpu.CANON_DEVICE))
    with pytest.raises(gpu.BtcError):
        gpu.Graph.from_coo(np.array([0, -1], np.int32), np.array([1, 2], np.int32), 5, 3 | gpu.CANON_DEVICE)
    with pytest.raises(gpu.BtcError):
        gpu.Graph.from_coo(np.array([0, 9], np.int32), np.array([1, 2], np.int32), 5, 3 | gpu.CANON_DEVICE)


def test_device_canon_from_device_pointers(gpu, orc):
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")  # the HIP runtime the library itself uses
    s, d = gpu.gen_rmat(14, 16, seed=3)
    ps, pd = C.c_void_p(), C.c_void_p()
    nb = C.c_size_t(s.nbytes)
    assert hip.hipMalloc(C.byref(ps), nb) == 0 and hip.hipMalloc(C.byref(pd), nb) == 0
    try:
        assert hip.hipMemcpy(ps, s.ctypes.data_as(C.c_void_p), nb, 1) == 0  # host to device
        assert hip.hipMemcpy(pd, d.ctypes.data_as(C.c_void_p), nb, 1) == 0
        dev = gpu.Graph.from_coo_device(ps.value, pd.value, s.size, 1 << 14)
    finally:
        hip.hipFree(ps)
        hip.hipFree(pd)
    host = gpu.Graph.from_coo(s, d, 1 << 14)
    same_graph(host, dev)
    want, _ = oracle_counts(orc, host)
    assert gpu.Plan(dev).count()[0] == want


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("K,tau", [(64, -1), (32, 0), (128, 0), (64, 100000)])
def test_device_planner_matches_oracle(gpu, orc, name, K, tau):
    s, d, n = case(name, gpu)
    g = gpu.Graph.from_coo(s, d, n)
    want, want_pv = oracle_counts(orc, g)
    p = gpu.Plan(g, tile_k=K, tau=tau, planner=2)
    t, pv, _ = p.count(per_vertex=True)
    assert t == want
    np.testing.assert_array_equal(pv, want_pv)
    hi = gpu.Plan(g, tile_k=K, tau=tau, planner=1).info
    di = p.info
    for k in PLAN_KEYS:
        assert di[k] == hi[k], k


@pytest.mark.parametrize("chunk", [1, 64, 4096])
def test_device_planner_item_size_invariant(gpu, orc, chunk):
    s, d, n = case("rmat15", gpu)
    g = gpu.Graph.from_coo(s, d, n, 3 | gpu.CANON_DEVICE)
    want, want_pv = oracle_counts(orc, g)
    for lookup in (0, 1, 2):
        t, pv, _ = gpu.Plan(g, chunk_work=chunk, lookup=lookup, planner=2).count(per_vertex=True)
        assert t == want
        np.testing.assert_array_equal(pv, want_pv)


def test_device_path_rmat20(gpu):
    """Device canonicalisation + device planner = host path, R-MAT s20."""
    s, d = gpu.gen_rmat(20, 16, seed=20)
    host = gpu.Graph.from_coo(s, d, 1 << 20)
    dev = gpu.Graph.from_coo(s, d, 1 << 20, 3 | gpu.CANON_DEVICE)
    for k in INFO_KEYS:
        assert host.info[k] == dev.info[k], k
    t_host, pv_host, _ = gpu.Plan(host, planner=1).count(per_vertex=True)
    t_dev, pv_dev, _ = gpu.Plan(dev, planner=2).count(per_vertex=True)
    assert t_host == t_dev
    np.testing.assert_array_equal(pv_host, pv_dev)
    assert int(pv_dev.sum()) == t_dev
    same_graph(host, dev)
