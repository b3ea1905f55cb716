"""Autotuner table (CPU) and search (GPU)."""
import numpy as np
import pytest


def bits(x):
    return 0 if x <= 1 else (x - 1).bit_length()


def test_table_roundtrip_and_lookup(btc, tmp_path):
    btc.tune_table_clear()
    s, d = btc.gen_rmat(12, 16, seed=5)
    g = btc.Graph.from_coo(s, d, 1 << 12)
    i = g.info
    key = (bits(i["n"]), bits(i["m"]), bits(i["wedges"]), bits(i["max_row"]))
    assert btc.tune_lookup(g) is None
    p = tmp_path / "tune.txt"
    p.write_text("btc-tune-3 %d %d %d %d 128 700 16 2 6 2 1024 1.2345\n" % key)
    btc.tune_table_load(p)
    o = btc.tune_lookup(g)
    assert (o["tile_k"], o["tau"], o["group_lanes"], o["unroll"], o["waves_per_eu"], o["dense_kernel"],
            o["core_ids"]) == (128, 700, 16, 2, 6, 2, 1024)
    q = tmp_path / "tune2.txt"
    btc.tune_table_save(q)
    assert q.read_text().split()[:12] == p.read_text().split()[:12]
    # four lanes per list (the default shape) is a valid entry
    p4 = tmp_path / "tune4.txt"
    p4.write_text("btc-tune-3 %d %d %d %d 64 128 4 1 0 1 2048 0.5\n" % key)
    btc.tune_table_load(p4)
    assert btc.tune_lookup(g)["group_lanes"] == 4
    bad = tmp_path / "bad.txt"
    bad.write_text("btc-tune-3 1 2 3 4 96 7 8 1 0 1 0 1.0\n")  # K = 96 is not a tile width
    with pytest.raises(btc.BtcError):
        btc.tune_table_load(bad)
    bad.write_text("btc-tune-3 1 2 3 4 64 7 2 1 0 1 0 1.0\n")  # 2 lanes per list is not a shape
    with pytest.raises(btc.BtcError):
        btc.tune_table_load(bad)
    btc.tune_table_clear()
    assert btc.tune_lookup(g) is None


@pytest.mark.gpu
def test_autotune_picks_exact_options(gpu, orc):
    gpu.tune_table_clear()
    s, d = gpu.gen_planted(8192, 128, 0.5, 6.0, seed=4)
    g = gpu.Graph.from_coo(s, d, 8192)
    rp, c = g.csr()
    want = orc.tc_merge(rp, c, threads=4)
    best, ms = gpu.autotune(g, budget_ms=3000, reps=2)
    assert best["tile_k"] in (32, 64, 128) and ms > 0
    assert gpu.tune_lookup(g)["tile_k"] == best["tile_k"]
    # a plan left on auto now takes the tuned options, and stays exact
    p = gpu.Plan(g)
    assert p.info["tile_k"] == best["tile_k"] and p.info["tau"] == best["tau"]
    assert p.count()[0] == want
    gpu.tune_table_clear()
