"""Exact triangle counting throughput on one MI355X.

Default workload: R-MAT scale 22, edge factor 16, (a, b, c, d) = (0.57, 0.19,
0.19, 0.05), seed 1, symmetrised, deduplicated, self-loops removed, ids
compacted and relabelled by descending degree.  One step = one exact count
of the whole graph by the hybrid engine (dense-tile kernel + wedge kernel)
with the graph and its work lists already resident in HBM; every step's
count is checked against the first and, once, against the CPU baseline.

Prints progress to stderr and ONE JSON line to stdout.

  python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--workload rmat22]
"""
import argparse
import json

import numpy as np
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "btc-tc-artifact_amd"))

METRIC = "exact-TC edges/s at R-MAT s22 (1 GPU); batched graphs/s + p99 at fixed QPS"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def make_workload(btc, name):
    if name.startswith("rmat"):
        scale = int(name[4:] or 22)
        src, dst = btc.gen_rmat(scale, 16, 0.57, 0.19, 0.19, seed=1)
        return src, dst, 1 << scale, {"workload": "rmat-s%d-ef16" % scale, "scale": scale, "edge_factor": 16,
                                       "abcd": [0.57, 0.19, 0.19, 0.05], "seed": 1}
    if name == "er":
        src, dst = btc.gen_er(16384, 16.0, seed=1)
        return src, dst, 16384, {"workload": "er-n16384-d16", "n": 16384, "avg_degree": 16, "seed": 1}
    if name == "stencil":
        src, dst = btc.gen_stencil3d(64, 64, 64)
        return src, dst, 64 ** 3, {"workload": "stencil27-64x64x64"}
    if name == "planted":
        src, dst = btc.gen_planted(1 << 17, 128, 0.5, 4.0, seed=1)
        return src, dst, 1 << 17, {"workload": "planted-2^17-c128-p0.5-noise4", "seed": 1}
    raise SystemExit("unknown workload " + name)


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo) and logical CPU count."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_threads():
    """Threads of the CPU share this job was given: OMP_NUM_THREADS when set
    (one GPU's share of the host), else every logical CPU."""
    return int(os.environ.get("BTC_CPU_THREADS") or os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)


def cpu_baseline(g, gpu_total):
    """The reference's own CPU verifier (oracle/_ref) on the same canonical L,
    with the threads of this job's CPU share; the oracle's C port when the
    reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    threads = cpu_threads()
    model, nproc = cpu_model()
    rowptr, col = g.csr()
    t0 = time.perf_counter()
    if oracle.ref_available():
        kind = "reference"
        total = oracle.ref_cpu_tc(rowptr, col, threads)
    else:
        kind = "port"
        total = oracle.tc_merge(rowptr, col, threads)
    dt = time.perf_counter() - t0
    return {"value": g.m / dt, "unit": "edges/s", "cores": threads, "kind": kind,
            "sample": "whole graph, same degree-ordered L as the GPU, one run (%.1f s)" % dt,
            "seconds": dt, "total": int(total), "matches_gpu": int(total) == int(gpu_total),
            "cpu_model": model, "host_logical_cpus": nproc,
            "threads_from": "OMP_NUM_THREADS (this job's CPU share)" if os.environ.get("OMP_NUM_THREADS")
            else "all logical CPUs"}


def load_traffic(workload):
    """HBM bytes per launch of the wedge kernel from the committed rocprofv3
    counter summary (L2-miss read requests by size, 128 / 64 / 32 bytes, plus
    WRITE_SIZE; checked on a device copy of known size); None when no summary
    exists for this workload."""
    path = os.path.join(ROOT, "profiles", "wedge_kernel_%s_traffic.md" % workload)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("hbm_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def copy_calibration(mb):
    """One device-to-device copy of `mb` MiB (HIP's blit copy kernel), so a
    rocprofv3 counter run can check its byte formulas on a kernel of known
    traffic."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mb << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log("copy calibration: %d MiB device-to-device" % mb)


VALU_BITPAIRS_PEAK = 6.3e14   # 256 CUs x 2.4 GHz x 64 lanes: 2 v_and_b32 + 2 v_bcnt per 64 bit pairs
MFMA_I8_MAC_PEAK = 2.5e15     # 5.0 POP/s dense int8 = 2.5e15 multiply-accumulates/s


def dense_ops(b, tau):
    """Algorithmic work of the dense path over Bit-BSR b: for every mask tile
    with popcount > tau, every slab (row block Jo) of its 16 x K columns and
    every tile column matched between its row block and Jo, 16 * 16 * K bit
    pairs.  Returns (dense tiles, matched pairs, bit pairs)."""
    indptr, indices, rows, tiles = b.to_host()
    K = b.view["tile_k"]
    wpt = K // 4
    pop = np.unpackbits(tiles.view(np.uint8).reshape(-1, wpt * 8), axis=1).sum(axis=1)
    dense = np.nonzero(pop > tau)[0]
    n_rb = len(indptr) - 1
    cols = {}

    def tc(I):
        if I not in cols:
            cols[I] = set(indices[indptr[I]:indptr[I + 1]].tolist())
        return cols[I]

    pairs = 0
    for t in dense.tolist():
        A = tc(int(rows[t]))
        for s in range(K // 16):
            Jo = int(indices[t]) * (K // 16) + s
            if Jo < n_rb:
                pairs += len(A & tc(Jo))
    return len(dense), pairs, pairs * 256 * K


def dense_leg(btc, runs):
    """Dense tiles, AND + popcount against i8 MFMA: the planted-community
    graph (2^17 vertices, communities of 128 with p = 0.5, ER noise of
    average degree 4) and the 27-point stencil 64^3, both in natural
    (compacted input) order so that dense tiles survive."""
    out = {}
    work = [("planted-2^17-c128-p0.5-noise4", lambda: (btc.gen_planted(1 << 17, 128, 0.5, 4.0, seed=1), 1 << 17)),
            ("stencil27-64x64x64", lambda: (btc.gen_stencil3d(64, 64, 64), 64 ** 3))]
    for name, make in work:
        (s, d), n = make()
        g = btc.Graph.from_coo(s, d, n, btc.CANON_COMPACT)
        want = None
        res = {"m": g.m}
        for K in (32, 64, 128):
            b = btc.BitBSR(g, K)
            tau = btc.default_tau(K)
            nd, pairs, ops = dense_ops(b, tau)
            row = {"tau": tau, "dense_tiles": nd, "matched_pairs": pairs, "bit_pairs": ops}
            for mf in (False, True):
                b.count(tau=tau, mfma=mf)
                ms = []
                for _ in range(runs):
                    t, _, k = b.count(tau=tau, mfma=mf)
                    want = t if want is None else want
                    assert t == want, (name, K, mf, t, want)
                    ms.append(k)
                row["mfma_ms" if mf else "and_popc_ms"] = round(statistics.median(ms), 4)
            # engine: dense-tile kernel alone (sparse edges go to the wedge kernel)
            p = btc.Plan(g, tile_k=K, tau=tau)
            for dk in (1, 2):
                p.tune(dense_kernel=dk)
                p.count(per_vertex=False)
                dms = []
                for _ in range(runs):
                    t, _, _ = p.count()
                    assert t == want, (name, K, dk)
                    dms.append(p.last_timing()[0])
                row["engine_dense_%s_ms" % ("mfma" if dk == 2 else "and_popc")] = round(statistics.median(dms), 4)
            if ops and row["engine_dense_and_popc_ms"] > 0:
                a = ops / (row["engine_dense_and_popc_ms"] * 1e-3)
                f = ops / (row["engine_dense_mfma_ms"] * 1e-3)
                row["and_popc_bitpairs_per_s"] = a
                row["and_popc_frac_valu"] = round(a / VALU_BITPAIRS_PEAK, 4)
                row["mfma_macs_per_s"] = f
                row["mfma_frac_i8"] = round(f / MFMA_I8_MAC_PEAK, 4)
            res["K%d" % K] = row
            p.close()
            b.close()
        res["triangles"] = int(want)
        log("dense %s: %s" % (name, res))
        out[name] = res
    return out


def configs_leg(btc, runs):
    """The other single-graph configurations: Erdos-Renyi n = 16384 with
    average degree 16, the 27-point stencil 64^3 and the planted-community
    graph 2^17.  Each: the engine (degree-ordered L, plan resident), the
    reference-semantics kernels at the Lite tile width on the natural
    (compacted input) order, and the library's host merge count
    (btc_count_cpu); all three totals must agree (and the stencil's must
    equal its closed form)."""
    out = {}
    for name in ("er", "stencil", "planted"):
        src, dst, n_in, cfg = make_workload(btc, name)
        g = btc.Graph.from_coo(src, dst, n_in, btc.CANON_DEFAULT | btc.CANON_DEVICE)
        p = btc.Plan(g)
        p.count()
        ks = []
        for _ in range(runs):
            t, _, ms = p.count()
            ks.append(ms)
        gn = btc.Graph.from_coo(src, dst, n_in, btc.CANON_COMPACT)
        K = btc.lite_tile_k(gn.n, gn.m)
        b = btc.BitBSR(gn, K)
        b.count()
        pk = []
        for _ in range(runs):
            tp, _, ms = b.count()
            pk.append(ms)
        threads = cpu_threads()
        btc.count_cpu(g, threads)  # the host copy of L is made on first use
        t0 = time.perf_counter()
        tc, _ = btc.count_cpu(g, threads)
        cpu_s = time.perf_counter() - t0
        exact = t == tp == int(tc) and (name != "stencil" or t == 11049696)
        km, pm = statistics.median(ks), statistics.median(pk)
        out[cfg["workload"]] = {"n": g.n, "m": g.m, "triangles": int(t), "exact": exact,
                                "engine_ms": round(km, 4), "engine_edges_per_s": g.m / (km * 1e-3),
                                "plain_tile_k": K, "plain_ms": round(pm, 4), "plain_edges_per_s": gn.m / (pm * 1e-3),
                                "engine_speedup": round(pm / km, 2), "cpu_s": round(cpu_s, 4),
                                "cpu_edges_per_s": g.m / cpu_s, "cpu_threads": threads}
        log("config %s: %s" % (cfg["workload"], out[cfg["workload"]]))
        assert exact, ("counts differ", name, t, tp, tc)
        p.close()
        b.close()
    return out


def batched_leg(btc, count, steps, sample, tune_ms):
    """4096 graphs (40 % ER, 40 % R-MAT, 20 % 9-point stencils, n log-uniform
    in [256, 16384]) counted by one persistent launch with the options the
    autotuner picks for the batch, against one launch per graph on a sample
    of them (each with its plan resident)."""
    t0 = time.perf_counter()
    work = btc.batch_workload(count, seed=1)
    t_gen = time.perf_counter() - t0
    # the batch built on the GPU straight from the pairs (one segmented
    # canonicalisation + one plan), against per-graph host canonicalisation
    t0 = time.perf_counter()
    gb = btc.Batch.from_coo([(s, d, n) for s, d, n, _ in work])
    t_gpu_build = time.perf_counter() - t0
    _, per_gpu, _ = gb.count()
    gb.close()
    t0 = time.perf_counter()
    graphs = [btc.Graph.from_coo(s, d, n) for s, d, n, _ in work]
    del work
    t_canon = time.perf_counter() - t0
    t0 = time.perf_counter()
    batch = btc.Batch(graphs)
    t_build = time.perf_counter() - t0
    info = batch.info
    batch.count()
    gpu_build_exact = bool((batch.count()[1] == per_gpu).all())
    log("batch built on the GPU from pairs in %.2f s (generation %.1f s); per-graph counts equal: %s"
        % (t_gpu_build, t_gen, gpu_build_exact))
    wall = []
    per = None
    for _ in range(steps):
        t1 = time.perf_counter()
        total, per, ms = batch.count()
        wall.append(time.perf_counter() - t1)
    t_default = statistics.median(wall)
    log("batched, default options: %d graphs, %d edges, %.3f ms per launch (%.0f graphs/s); build %.1f s, canon %.1f s"
        % (count, info["m"], t_default * 1e3, count / t_default, t_build, t_canon))
    tuned = None
    if tune_ms > 0:
        t0 = time.perf_counter()
        tuned, tms = batch.autotune(budget_ms=tune_ms)
        t_tune = time.perf_counter() - t0
        wall = []
        for _ in range(steps):
            t1 = time.perf_counter()
            total_t, per_t, _ = batch.count()
            wall.append(time.perf_counter() - t1)
            assert total_t == total and bool((per_t == per).all()), "tuned batch count differs"
        log("batched, tuned %s in %.1f s: %.3f ms per launch" % (tuned, t_tune, statistics.median(wall) * 1e3))
        tuned = {k: tuned[k] for k in ("tile_k", "tau", "group_lanes", "unroll", "waves_per_eu", "dense_kernel")}
        tuned["tune_s"] = round(t_tune, 1)
    t_batch = statistics.median(wall)
    # one launch per graph, plans resident, on an evenly spaced sample
    idx = list(range(0, count, max(1, count // sample)))[:sample]
    plans = [btc.Plan(graphs[k]) for k in idx]
    for p in plans[:8]:
        p.count()
    t1 = time.perf_counter()
    single = [p.count()[0] for p in plans]
    t_single = (time.perf_counter() - t1) / len(plans)
    exact = all(int(per[k]) == single[j] for j, k in enumerate(idx))
    log("one launch per graph: %.3f ms per graph (%.0f graphs/s) on %d graphs; exact %s"
        % (t_single * 1e3, 1 / t_single, len(plans), exact))
    # the same with every count replayed from a captured HIP graph
    for p in plans:
        p.capture()
    t1 = time.perf_counter()
    replay = [p.count()[0] for p in plans]
    t_replay = (time.perf_counter() - t1) / len(plans)
    exact = exact and replay == single
    batch.capture()
    wall_r = []
    for _ in range(steps):
        t1 = time.perf_counter()
        total_r, per_r, _ = batch.count()
        wall_r.append(time.perf_counter() - t1)
    exact = exact and total_r == total and bool((per_r == per).all())
    log("replayed: %.3f ms per single graph, %.3f ms per batch launch" % (t_replay * 1e3,
                                                                         statistics.median(wall_r) * 1e3))
    return {"graphs": count, "edges": info["m"], "ms_per_launch": round(t_batch * 1e3, 3), "tuned_options": tuned,
            "default_ms_per_launch": round(t_default * 1e3, 3),
            "graphs_per_s": count / t_batch, "edges_per_s": info["m"] / t_batch,
            "single_launch_graphs_per_s": 1.0 / t_single, "speedup_vs_single": round(t_single * count / t_batch, 2),
            "single_sample": len(plans), "exact_vs_single": exact, "triangles": int(total),
            "single_replay_graphs_per_s": 1.0 / t_replay, "ms_per_launch_replay": round(statistics.median(wall_r) * 1e3, 3),
            "build_s": round(t_build, 2), "canon_s": round(t_canon, 2), "work_items": info["plan"]["work_items"],
            "gpu_build_from_pairs_s": round(t_gpu_build, 3), "gpu_build_exact": gpu_build_exact}


def serving_leg(btc, requests, seed=1, canon_flags=None, prep_threads=None, group_graphs=None, captured_prep=None):
    """Poisson stream of exact-count queries: R-MAT graphs (edge factor 16)
    with n log-normal (median 2^13, sigma 1 in ln n, capped at 2^20), 30 %
    of requests repeat a graph submitted earlier.  A closed burst measures
    saturation; the open-loop run uses 70 % of it.  Latency = done - submit
    on the server clock."""
    import threading

    rng = np.random.default_rng(seed)
    draws = iter(range(seed * 1000, seed * 1000 + 1000))

    def stream(nreq, key0):
        return btc.serving_workload(nreq, seed=next(draws), key0=key0)

    warm = stream(48, 3 << 40)

    def run(reqs, qps):
        sv = btc.Server(canon_flags=canon_flags, prep_threads=prep_threads, group_graphs=group_graphs,
                        captured_prep=captured_prep)
        for key, g in warm:
            if g is not None:
                sv.submit(key, g[0], g[1], g[2])
        done = 0
        while done < sum(1 for _, g in warm if g is not None):
            done += len(sv.poll(wait_ms=20))
        gaps = rng.exponential(1.0 / qps, len(reqs)) if qps else np.zeros(len(reqs))
        t_start = time.perf_counter()

        def submit():
            t = 0.0
            for (key, g), gap in zip(reqs, gaps):
                t += gap
                delay = t - (time.perf_counter() - t_start)
                if delay > 0:
                    time.sleep(delay)
                if g is None:
                    sv.submit(key)
                else:
                    sv.submit(key, g[0], g[1], g[2])

        th = threading.Thread(target=submit)
        th.start()
        res = []
        while len(res) < len(reqs):
            res += sv.poll(wait_ms=20)
        th.join()
        st = sv.stats()
        sv.close()
        return res, st

    # a closed burst measures saturation; each run first warms its server
    # (preparation arenas and pinned staging grow on first use) with graphs
    # of other keys
    burst = stream(max(64, 2 * requests // 3), 1 << 40)
    res, _ = run(burst, 0)
    span = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
    sat = len(res) / span
    qps = 0.7 * sat
    reqs = stream(requests, 2 << 40)
    res, st = run(reqs, qps)
    lat = np.array(sorted(r["t_done"] - r["t_submit"] for r in res))
    wait = np.array([r["t_start"] - r["t_submit"] for r in res if not r["resident_hit"]])
    run_ = np.array([r["t_done"] - r["t_start"] for r in res])
    span = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
    graphs = {}
    for key, g in reqs:
        if g is not None:
            graphs[key] = g
    want = {k: btc.Plan(btc.Graph.from_coo(*g)).count()[0] for k, g in graphs.items()}
    exact = all(r["status"] == 0 and r["triangles"] == want[r["graph_key"]] for r in res)
    out = {"requests": len(res), "qps": round(qps, 1), "saturation_graphs_per_s": round(sat, 1),
           "graphs_per_s": round(len(res) / span, 1), "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
           "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3), "exact_alone_vs_mix": exact,
           "resident_hits": st["resident_hits"], "launches": st["launches"],
           "mean_batch": round(st["launched_requests"] / max(1, st["launches"]), 2),
           "distinct_graphs": len(graphs),
           "p50_prepare_ms": round(1e3 * float(np.percentile(wait, 50)), 3) if wait.size else None,
           "p50_launch_ms": round(1e3 * float(np.percentile(run_, 50)), 3)}
    log("serving (canon flags %s, captured preparation %s): %s" % (canon_flags, captured_prep, out))
    return out


def pipeline_leg(btc, graphs, seed=3):
    """Captured build + count (btc_pipeline) against the eager device path on
    `graphs` new graphs of the serving workload: one pipeline per size class
    (ids and pairs rounded up to powers of two), every graph run through it
    twice (the second run timed: wall from host pairs to total, and the
    replay by HIP events) and built eagerly (btc_graph_from_coo_device path +
    plan + count, wall); every total equal to the eager one."""
    work = [g for _, g in btc.serving_workload(3 * graphs, seed=seed) if g is not None][:graphs]

    def cls(n, nnz):
        return 1 << max(10, (n - 1).bit_length()), 1 << max(12, (max(nnz, 1) - 1).bit_length())

    pipes = {}
    t0 = time.perf_counter()
    for s, d, n in work:
        c = cls(n, s.size)
        if c not in pipes:
            pipes[c] = btc.Pipeline(c[0], c[1], btc.CANON_DEFAULT)
    create_s = time.perf_counter() - t0
    eager, wall, replay, exact = [], [], [], True
    for s, d, n in work:
        p = pipes[cls(n, s.size)]
        p.count(s, d, n)
        t1 = time.perf_counter()
        got, ms = p.count(s, d, n)
        t2 = time.perf_counter()
        want = btc.Plan(btc.Graph.from_coo(s, d, n, btc.CANON_DEFAULT | btc.CANON_DEVICE)).count()[0]
        t3 = time.perf_counter()
        exact = exact and got == want
        wall.append(t2 - t1)
        eager.append(t3 - t2)
        replay.append(ms)
    nodes = sorted({p.info["graph_nodes"] for p in pipes.values()})
    out = {"graphs": len(work), "size_classes": len(pipes), "graph_nodes": nodes,
           "create_s": round(create_s, 3), "median_pairs": int(np.median([g[0].size for g in work])),
           "pipeline_wall_ms_p50": round(1e3 * float(np.median(wall)), 3),
           "replay_ms_p50": round(float(np.median(replay)), 3),
           "eager_wall_ms_p50": round(1e3 * float(np.median(eager)), 3),
           "speedup_p50": round(float(np.median(eager)) / float(np.median(wall)), 2), "exact": exact}
    log("pipeline: %s" % out)
    for p in pipes.values():
        p.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="rmat22")
    ap.add_argument("--tile-k", type=int, default=0)
    ap.add_argument("--tau", type=int, default=-1)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--plain-runs", type=int, default=3, help="reference-semantics Bit-BSR yardstick runs (0 = skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch", type=int, default=4096, help="graphs in the batched leg (0 = skip)")
    ap.add_argument("--batch-sample", type=int, default=256, help="graphs timed one launch each")
    ap.add_argument("--batch-tune-ms", type=int, default=3000, help="autotuner budget for the batch (0 = defaults)")
    ap.add_argument("--serve", type=int, default=600, help="requests in the serving leg (0 = skip)")
    ap.add_argument("--dense", type=int, default=5, help="timed runs per case in the dense-tile leg (0 = skip)")
    ap.add_argument("--host-path", type=int, default=1, help="also time host canonicalisation and planning")
    ap.add_argument("--configs", type=int, default=5, help="timed runs per case of the ER / stencil / planted leg")
    ap.add_argument("--shares", type=int, default=1, help="time one GPU's share of 2-, 4- and 8-way splits")
    ap.add_argument("--serve-host-prep", type=int, default=1,
                    help="also run the serving leg with canonicalisation and planning on host threads")
    ap.add_argument("--serve-eager-prep", type=int, default=1,
                    help="also run the serving leg with eager (uncaptured, grouped) GPU preparation")
    ap.add_argument("--serve-threads", type=int, default=0, help="prep threads of the serving legs (0 = default)")
    ap.add_argument("--pipeline", type=int, default=96,
                    help="serving-workload graphs run through captured build + count pipelines (0 = skip)")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="run one D2D copy of this many MiB first (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the engine has no multi-GPU path "
              "(a k-way split would give each GPU a 1/k share of the work items and all-reduce one u64); "
              "run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)

    import btc_amd as btc

    rc = btc.lib().btc_device_init(0)
    if rc != 0:
        raise SystemExit("no HIP device: " + btc.lib().btc_last_error().decode())

    if args.copy_calibration_mb > 0:
        copy_calibration(args.copy_calibration_mb)
    t0 = time.perf_counter()
    src, dst, n_in, config = make_workload(btc, args.workload)
    log("generated %d pairs in %.2f s" % (src.size, time.perf_counter() - t0))
    # end to end from host pairs: GPU canonicalisation, GPU planner, first count
    t0 = time.perf_counter()
    g = btc.Graph.from_coo(src, dst, n_in, btc.CANON_DEFAULT | btc.CANON_DEVICE)
    canon_ms = (time.perf_counter() - t0) * 1e3
    gi = g.info
    log("canonical L: n=%d m=%d max_row=%d wedges=%d (%.0f ms on the GPU)"
        % (gi["n"], gi["m"], gi["max_row"], gi["wedges"], canon_ms))
    t0 = time.perf_counter()
    plan = btc.Plan(g, tile_k=args.tile_k, tau=args.tau, chunk_work=args.chunk)
    build_ms = (time.perf_counter() - t0) * 1e3
    pi = plan.info
    log("plan: %s (%.0f ms)" % (pi, build_ms))
    t0 = time.perf_counter()
    plan.count()
    first_ms = (time.perf_counter() - t0) * 1e3
    e2e = {"canon_ms": round(canon_ms, 1), "plan_ms": round(build_ms, 1), "first_count_ms": round(first_ms, 2),
           "e2e_ms": round(canon_ms + build_ms + first_ms, 1),
           "e2e_edges_per_s": gi["m"] / ((canon_ms + build_ms + first_ms) * 1e-3)}
    # the same three steps again on the same pairs with new objects: the
    # first pass above also pays the process's one-time costs (kernel code
    # loading, the pinned staging buffer, the device memory pool's growth)
    t0 = time.perf_counter()
    g2 = btc.Graph.from_coo(src, dst, n_in, btc.CANON_DEFAULT | btc.CANON_DEVICE)
    t1 = time.perf_counter()
    p2 = btc.Plan(g2, tile_k=args.tile_k, tau=args.tau, chunk_work=args.chunk)
    t2 = time.perf_counter()
    warm_total = p2.count()[0]
    t3 = time.perf_counter()
    p2.close()
    g2.close()
    del p2, g2
    e2e.update({"warm_canon_ms": round((t1 - t0) * 1e3, 1), "warm_plan_ms": round((t2 - t1) * 1e3, 1),
                "warm_first_count_ms": round((t3 - t2) * 1e3, 2), "warm_e2e_ms": round((t3 - t0) * 1e3, 1),
                "warm_e2e_edges_per_s": gi["m"] / (t3 - t0), "warm_count_equal": warm_total == plan.count()[0]})
    if args.host_path:
        # the same steps on host threads, for comparison
        t0 = time.perf_counter()
        gh = btc.Graph.from_coo(src, dst, n_in)
        e2e["host_canon_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        t0 = time.perf_counter()
        ph = btc.Plan(gh, tile_k=args.tile_k, tau=args.tau, chunk_work=args.chunk, planner=1)
        e2e["host_plan_ms"] = round((time.perf_counter() - t0) * 1e3, 1)
        hi = ph.info
        e2e["host_plan_same_stats"] = all(hi[k] == pi[k] for k in ("sparse_edges", "dense_edges", "dense_tiles",
                                                                   "wedge_checks", "core_lists", "core_rows"))
        e2e["host_count_equal"] = ph.count()[0] == plan.count()[0]
        ph.close()
        gh.close()
        del ph, gh
    log("end to end: %s" % e2e)
    del src, dst

    total0 = None
    for _ in range(args.warmup):
        total0, _, _ = plan.count()
    kms, wms, totals = [], [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t, _, ms = plan.count()
        totals.append(t)
        kms.append(ms)
        wms.append(plan.last_timing()[1])
    wall = time.perf_counter() - t0
    if total0 is None:
        total0 = totals[0]
    assert all(t == total0 for t in totals), "count changed between steps"
    ms_per_step = wall * 1e3 / args.steps
    value = gi["m"] / (wall / args.steps)
    log("engine: %d triangles, %.3f ms/step wall, kernel median %.3f ms, wedge kernel mean %.3f ms"
        % (total0, ms_per_step, statistics.median(kms), statistics.mean(wms)))

    # roofline of the dominant kernel (wedge kernel): the bytes this
    # algorithm must stream per launch (plan.stream_bytes: looked-up prefix
    # ids at 2 or 4 B, pivot rows and hub-core words at 4 B, 16 B per list
    # record, 32 B per work item) over its mean HIP-event time on the launch
    # stream.  The merge-intersection stream of the same edges, B = 4(n+1) +
    # 4 m_s + 4 W_s, is kept as context: this kernel does not read it.
    wedge_ms = statistics.mean(wms)
    b_merge = 4 * (gi["n"] + 1) + 4 * pi["sparse_edges"] + 4 * pi["sparse_merge_work"]
    stream_b = pi["stream_bytes"]
    achieved = stream_b / (wedge_ms * 1e-3) / 1e9 if wedge_ms > 0 else 0.0
    traffic = load_traffic(config["workload"])
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_GBps": round(traffic / (wedge_ms * 1e-3) / 1e9, 1) if traffic and wedge_ms > 0 else None,
                "traffic_frac": round(traffic / (wedge_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                if traffic and wedge_ms > 0 else None,
                "traffic_over_alg": round(traffic / stream_b, 3) if traffic and stream_b else None,
                "kernel": "wedge_kernel", "kernel_ms": round(wedge_ms, 4), "alg_bytes": stream_b,
                "alg_bytes_model": "plan.stream_bytes: 2 or 4 B per looked-up prefix id, 4 B per pivot-row id and "
                                   "hub-core word, 16 B per list record, 32 B per work item",
                "merge_stream_bytes": b_merge,
                "merge_stream_GBps": round(b_merge / (wedge_ms * 1e-3) / 1e9, 1) if wedge_ms > 0 else None}
    if roofline["frac"] > 1.0:
        log("warning: the kernel's byte stream exceeds the HBM peak (cache reuse): %s" % roofline)

    # one GPU's share of a k-way split (the work each GPU would run), each
    # share on the whole chip, from a plan whose work items are cut for a
    # 1/k share (item size = the single-GPU rule divided by k); the shares'
    # totals must add up
    sharding = None
    if args.shares:
        sharding = {}
        full_ms = statistics.median(kms)
        cus = 256
        for k in (2, 4, 8):
            chunk = int(max(1024, min(1 << 18, gi["wedges"] // (cus * 64 * k))))
            pk = btc.Plan(g, tile_k=args.tile_k, tau=args.tau, chunk_work=chunk)
            ms_k, tot = [], 0
            for j in range(k):
                t_j, ms = pk.count_share(k, j)
                ms_j = [pk.count_share(k, j)[1] for _ in range(3)]
                ms_k.append(statistics.median(ms_j))
                tot += t_j
            pk.close()
            sharding[str(k)] = {"chunk_work": chunk, "max_share_ms": round(max(ms_k), 4),
                                "min_share_ms": round(min(ms_k), 4),
                                "projected_speedup": round(full_ms / max(ms_k), 2), "shares_add_up": tot == total0}
            assert tot == total0, ("shares do not add up", k, tot, total0)
        log("sharding (one GPU's share of a k-way split): %s" % sharding)

    plain = None
    if args.plain_runs > 0:
        K = 128 if gi["n"] >= 50000 else btc.lite_tile_k(gi["n"], gi["m"])
        t0 = time.perf_counter()
        b = btc.BitBSR(g, K)
        bms = (time.perf_counter() - t0) * 1e3
        b.count()  # warmup
        pms = []
        for _ in range(args.plain_runs):
            t, _, ms = b.count()
            assert t == total0, ("reference-semantics path disagrees", t, total0)
            pms.append(ms)
        pmed = statistics.median(pms)
        plain = {"tile_k": K, "kernel_ms": round(pmed, 3), "edges_per_s": gi["m"] / (pmed * 1e-3),
                 "build_ms": round(bms, 1), "num_blocks": b.view["num_blocks"],
                 "engine_speedup": round(pmed / statistics.median(kms), 2)}
        log("plain 16x%d: %.3f ms (engine speedup %.2fx)" % (K, pmed, plain["engine_speedup"]))
        b.close()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(g, total0)
        log("cpu baseline: %s" % cpu)
        assert cpu["matches_gpu"], "GPU count differs from the CPU baseline"

    batched = (batched_leg(btc, args.batch, max(3, args.steps // 2), args.batch_sample, args.batch_tune_ms)
               if args.batch > 0 else None)
    pt = args.serve_threads or None
    serving = serving_leg(btc, args.serve, prep_threads=pt) if args.serve > 0 else None
    serving_eager = (serving_leg(btc, args.serve, prep_threads=pt, captured_prep=0)
                     if args.serve > 0 and args.serve_eager_prep else None)
    serving_host = (serving_leg(btc, args.serve, canon_flags=btc.CANON_DEFAULT, prep_threads=pt)
                    if args.serve > 0 and args.serve_host_prep else None)
    pipeline = pipeline_leg(btc, args.pipeline) if args.pipeline > 0 else None
    dense = dense_leg(btc, args.dense) if args.dense > 0 else None
    configs = configs_leg(btc, args.configs) if args.configs > 0 else None

    config.update({"n": gi["n"], "m": gi["m"], "ordering": "degree-descending", "tile_k": pi["tile_k"],
                   "tau": pi["tau"], "dense_tiles": pi["dense_tiles"], "work_items": pi["work_items"]})
    out = {
        "metric": METRIC, "value": value, "unit": "edges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic", "config": config,
        "triangles": int(total0), "triangles_per_s": total0 / (wall / args.steps),
        "kernel_ms_median": statistics.median(kms), "canon_ms": round(canon_ms, 1), "plan_ms": round(build_ms, 1),
        "roofline": roofline, "cpu_baseline": cpu, "plain_hybrid": plain, "end_to_end": e2e, "sharding": sharding,
        "batched": batched, "serving": serving, "serving_eager_prep": serving_eager, "pipeline": pipeline,
        "serving_host_prep": serving_host, "dense_tiles": dense,
        "other_configs": configs,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
