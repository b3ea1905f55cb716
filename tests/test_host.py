"""Host-side library checks that need no GPU: the C ABI loads and exports
every symbol of include/btc_amd.h, the generators are deterministic and
shaped as specified, the host canonicalisation equals the oracle's bit for
bit, and argument errors come back as status codes."""
import os
import re

import numpy as np
import pytest

import graphs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "btc_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(btc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol(btc):
    names = header_functions()
    assert len(names) >= 28
    missing = [n for n in names if not hasattr(btc.lib(), n)]
    assert not missing, missing


def test_version(btc):
    assert btc.version().startswith("0.1.0")


def test_rmat_generator(btc):
    s, d = btc.gen_rmat(12, 8, seed=9)
    assert s.size == 8 << 12 and s.min() >= 0 and s.max() < 4096 and d.max() < 4096
    s2, d2 = btc.gen_rmat(12, 8, seed=9)
    np.testing.assert_array_equal(s, s2)
    np.testing.assert_array_equal(d, d2)
    s3, _ = btc.gen_rmat(12, 8, seed=10)
    assert not np.array_equal(s, s3)
    # a = 0.57: bit l of the source is 1 with probability c + d = 0.24
    frac = np.mean((s >> 3) & 1)
    assert abs(frac - 0.24) < 0.01


def test_rmat_thread_count_independent(btc):
    import subprocess, sys
    code = ("import sys; sys.path.insert(0, %r); import btc_amd, numpy as np; s, d = btc_amd.gen_rmat(14, 4, seed=3);"
            "print(int(np.bitwise_xor.reduce(s * 7919 + d)))") % os.path.join(ROOT, "btc-tc-artifact_amd")
    outs = {subprocess.run([sys.executable, "-c", code], env=dict(os.environ, OMP_NUM_THREADS=str(t)),
                           capture_output=True, text=True, check=True).stdout for t in (1, 3)}
    assert len(outs) == 1


def test_er_generator(btc):
    n, deg = 16384, 16.0
    s, d = btc.gen_er(n, deg, seed=1)
    assert np.all(s > d)  # pairs (i, j), j < i, no duplicates
    assert np.unique(s.astype(np.int64) * n + d).size == s.size
    mean_deg = 2 * s.size / n
    assert abs(mean_deg - deg) < 0.2  # sd of the mean is about 0.03


def test_stencil_generators(btc):
    s, d = btc.gen_stencil2d(7, 5)
    assert s.size == 6 * 5 + 7 * 4 + 2 * 6 * 4
    s, d = btc.gen_stencil3d(64, 64, 64)
    # 27-point: 3,298,428 undirected edges at 64^3
    assert s.size == 3298428


def test_planted_generator(btc):
    s, d = btc.gen_planted(1024, 128, 0.5, 0.0, seed=2)
    assert np.all(s // 128 == d // 128)  # no noise: only intra-community pairs
    assert abs(s.size / (8 * 128 * 127 / 2) - 0.5) < 0.02


@pytest.mark.parametrize("flags", [0, 1, 2, 3])
@pytest.mark.parametrize("seed", range(3))
def test_canon_matches_oracle(btc, orc, flags, seed):
    s, d = graphs.messy(3000, 40000, seed)
    g = btc.Graph.from_coo(s, d, 3100, flags)
    rp, c = g.csr()
    rp2, c2, lb2 = orc.canon(s, d, 3100, flags)
    np.testing.assert_array_equal(rp, rp2)
    np.testing.assert_array_equal(c, c2)
    np.testing.assert_array_equal(g.labels(), lb2)


def test_canon_rmat_matches_oracle(btc, orc):
    s, d = btc.gen_rmat(14, 16, seed=4)
    g = btc.Graph.from_coo(s, d, 1 << 14)
    rp, c = g.csr()
    rp2, c2, _ = orc.canon(s, d, 1 << 14, 3)
    np.testing.assert_array_equal(rp, rp2)
    np.testing.assert_array_equal(c, c2)
    info = g.info
    deg = np.diff(rp)
    assert info["max_row"] == deg.max() and info["wedges"] == int((deg * (deg - 1) // 2).sum())


def test_graph_from_csr_validation(btc):
    with pytest.raises(btc.BtcError) as e:
        btc.Graph.from_csr(np.array([0, 1], np.int64), np.array([0], np.int32))  # col == row
    assert e.value.code == 1
    with pytest.raises(btc.BtcError):
        btc.Graph.from_csr(np.array([0, 0, 2], np.int64), np.array([0, 0], np.int32))  # duplicate
    g = btc.Graph.from_csr(np.array([0, 0, 1, 3], np.int64), np.array([0, 0, 1], np.int32))
    assert g.m == 3 and btc.count_cpu(g)[0] == 1


def test_bad_input_is_a_status_not_an_exit(btc):
    with pytest.raises(btc.BtcError) as e:
        btc.Graph.from_coo(np.array([0, 9], np.int32), np.array([1, 2], np.int32), n=5)
    assert e.value.code == 1 and "vertex id" in str(e.value)
    with pytest.raises(btc.BtcError):
        btc.Graph.from_coo(np.array([-3], np.int32), np.array([1], np.int32))
    with pytest.raises(btc.BtcError):
        btc.Graph.from_coo(np.array([1], np.int32), np.array([0], np.int32), flags=8)


def test_empty_and_loop_only_graphs(btc):
    g = btc.Graph.from_coo(np.zeros(0, np.int32), np.zeros(0, np.int32), n=10)
    assert g.info["n"] == 10 and g.m == 0 and btc.count_cpu(g)[0] == 0
    g = btc.Graph.from_coo(np.array([3, 3], np.int32), np.array([3, 3], np.int32), n=5)
    assert g.m == 0


def test_count_cpu_matches_oracle(btc, orc):
    s, d = btc.gen_rmat(13, 16, seed=6)
    g = btc.Graph.from_coo(s, d, 1 << 13)
    rp, c = g.csr()
    t, pv = btc.count_cpu(g, per_vertex=True)
    t2, p2 = orc.tc_merge(rp, c, per_row=True)
    assert t == t2
    np.testing.assert_array_equal(pv, p2)


def test_lite_policy_matches_oracle(btc, orc):
    for n, m in [(16384, 131000), (262144, 3298428), (2395967, 64156180), (60000, 900000), (60000, 899999)]:
        for mode in (0, 1, 2):
            assert btc.lite_tile_k(n, m, mode) == orc.lite_tile_k(n, m, mode)
    assert [btc.default_tau(k) for k in (32, 64, 128)] == [64, 128, 512]


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with status 2 and says why."""
    import os
    import subprocess
    import sys

    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""
