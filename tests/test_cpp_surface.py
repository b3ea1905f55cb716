"""The C++ drop-in surface (btc/btc.h) and the Lite CLI.

CPU: a host program built against btc/btc.h reads every input format the
reference reads (.mtx general and symmetric, .el, .csr, .bin), preprocesses
and counts on the CPU verify path.  GPU: btc_tc_lite prints the reference's
stdout contract (apps/btc_tc_lite.cu:141-149) with the exact count on both
engines; the regular expressions are the contract's lines.
"""
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import graphs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "btc-tc-artifact_amd")
LIBDIR = os.path.join(PKG, "btc_amd", "lib")
LITE = os.path.join(PKG, "bin", "btc_tc_lite")


def write_inputs(tmp, s, d, n):
    """The same graph in the four formats of utils/io/read.h."""
    paths = {}
    p = os.path.join(tmp, "g.mtx")
    with open(p, "w") as f:
        f.write("%%MatrixMarket matrix coordinate pattern general\n% comment\n")
        f.write("%d %d %d\n" % (n, n, s.size))
        for a, b in zip(s, d):
            f.write("%d %d\n" % (a + 1, b + 1))
    paths["mtx"] = p
    # symmetric storage: lower triangle only, values present
    p = os.path.join(tmp, "gs.mtx")
    lo = {(max(a, b), min(a, b)) for a, b in zip(s.tolist(), d.tolist()) if a != b}
    with open(p, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n")
        f.write("%d %d %d\n" % (n, n, len(lo)))
        for a, b in sorted(lo):
            f.write("%d %d 1.5\n" % (a + 1, b + 1))
    paths["mtx_sym"] = p
    p = os.path.join(tmp, "g.el")
    with open(p, "w") as f:
        f.write("# Directed graph\n# Nodes: %d Edges: %d\n" % (n, s.size))
        for a, b in zip(s, d):
            f.write("%d\t%d\n" % (a, b))
    paths["el"] = p
    order = np.lexsort((d, s))
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, s[order] + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    col = d[order].astype(np.int32)
    p = os.path.join(tmp, "g.csr")
    with open(p, "wb") as f:
        f.write(struct.pack("<ii", n, s.size) + rp.tobytes() + col.tobytes())
    paths["csr"] = p
    p = os.path.join(tmp, "g.bin")
    with open(p, "wb") as f:
        f.write(struct.pack("<iii", n, n, s.size) + rp.tobytes() + col.tobytes() +
                np.ones(s.size, np.float32).tobytes())
    paths["bin"] = p
    return paths


@pytest.fixture(scope="module")
def header_check(btc, tmp_path_factory):
    out = str(tmp_path_factory.mktemp("cpp") / "header_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", PKG,
                    os.path.join(ROOT, "tests", "cpp", "header_check.cpp"), "-o", out, "-L", LIBDIR, "-lbtc_amd",
                    "-Wl,-rpath," + LIBDIR], check=True)
    return out


@pytest.mark.parametrize("fmt", ["mtx", "mtx_sym", "el", "csr", "bin"])
def test_header_reads_every_format(header_check, tmp_path, fmt):
    s, d, n, t = graphs.karate()
    paths = write_inputs(str(tmp_path), s, d, n)
    r = subprocess.run([header_check, paths[fmt]], capture_output=True, text=True, check=True)
    m = re.search(r"triangles (\d+) edges (\d+)", r.stdout)
    assert m and int(m.group(1)) == t == 45 and int(m.group(2)) == 78, r.stdout


def test_header_rmat_matches_oracle(btc, orc, header_check, tmp_path):
    s, d = btc.gen_rmat(11, 16, seed=3)
    paths = write_inputs(str(tmp_path), s, d, 1 << 11)
    rp, c, _ = orc.canon(s, d, 1 << 11, 3)
    want = orc.tc_merge(rp, c)
    for fmt in ("mtx", "el", "csr"):
        r = subprocess.run([header_check, paths[fmt]], capture_output=True, text=True, check=True)
        assert int(re.search(r"triangles (\d+)", r.stdout).group(1)) == want


def test_header_reports_bad_files(header_check, tmp_path):
    p = tmp_path / "bad.mtx"
    p.write_text("not a matrix\n")
    r = subprocess.run([header_check, str(p)], capture_output=True, text=True)
    assert r.returncode != 0 and "MatrixMarket" in r.stderr


def test_lite_cli_usage(btc):
    assert os.path.exists(LITE), "btc_tc_lite not built"
    r = subprocess.run([LITE], capture_output=True, text=True)
    assert r.returncode != 0 and "[-i input_file]" in r.stdout


CONTRACT = [r"\[Preprocessing\] time: [\d.]+ ms",
            r"\[Time Breakdown\] Convert: [\d.]+ ms, Compute \(Kernel\): [\d.]+ ms",
            r"\[Total Time \(Convert\+Compute\)\] time: [\d.]+ ms",
            r"Triangles \(GPU\): (\d+)",
            r"Kernel_runs:( [\d.]+)+"]


@pytest.mark.gpu
@pytest.mark.parametrize("engine,mode", [("hybrid", "0"), ("plain", "0"), ("plain", "2"), ("hybrid", "2")])
def test_lite_cli_contract(gpu, tmp_path, engine, mode):
    s, d, n, t = graphs.karate()
    paths = write_inputs(str(tmp_path), s, d, n)
    env = dict(os.environ, BTC_ENGINE=engine, BTC_MODE=mode)
    r = subprocess.run([LITE, "-i", paths["mtx"], "-w", "1", "-r", "3", "-v", "1"], capture_output=True,
                       text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    for pat in CONTRACT:
        assert re.search(pat, r.stdout), pat
    assert int(re.search(CONTRACT[3], r.stdout).group(1)) == t
    assert "Verify: PASS" in r.stdout
    assert ("USE 128" in r.stdout) == (mode == "2")


@pytest.fixture(scope="module")
def adaptive_check(btc, tmp_path_factory):
    out = str(tmp_path_factory.mktemp("cpp") / "adaptive_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", PKG,
                    os.path.join(ROOT, "tests", "cpp", "adaptive_check.cpp"), "-o", out, "-L", LIBDIR, "-lbtc_amd",
                    "-Wl,-rpath," + LIBDIR], check=True)
    return out


def test_adaptive_check_builds(adaptive_check):
    assert os.path.exists(adaptive_check)


@pytest.mark.gpu
@pytest.mark.parametrize("reorder", ["1", "0"])
def test_reference_entry_points_reach_engine(gpu, orc, adaptive_check, tmp_path, reorder):
    """count_triangles_16x{128,32}_adaptive (engine), v4/v5/v6 (reference
    semantics) on the same BCSR handle: every count equals the oracle's, in
    degree order and in the input's id order (BTC_REORDER=0)."""
    from test_gpu_long_rows import hub_graph
    cases = [gpu.gen_rmat(12, 16, seed=5) + (1 << 12,), hub_graph(seed=2, others=4000, degree=8)]
    if reorder == "0":
        cases.append(hub_graph(seed=3))  # an L-row of 66000 ids
    for s, d, n in cases:
        rp, c, _ = orc.canon(s, d, n, 3)
        want = orc.tc_merge(rp, c, threads=16)
        path = write_inputs(str(tmp_path), s, d, n)["el"]
        env = dict(os.environ, BTC_REORDER=reorder)
        r = subprocess.run([adaptive_check, path], capture_output=True, text=True, env=env, timeout=300)
        assert r.returncode == 0, r.stdout + r.stderr
        got = dict(re.findall(r"^(\w+) (\d+)$", r.stdout, re.M))
        calls = ["v4_16x128", "v5_16x128", "adaptive_16x128", "adaptive_16x128_again", "v5_16x32", "v6_16x32",
                 "adaptive_16x32", "adaptive_16x32_again"]
        for k in calls:
            assert int(got[k]) == want, (k, got, want)
        assert got["adaptive_16x128_ms_set"] == "1"
