"""ctypes binding of libbtc_amd.so, the gfx950 exact triangle-counting engine.

Everything here calls the C ABI declared in include/btc_amd.h; the counting
itself runs in the HIP kernels of that library.  If the library has not been
built, every call raises BtcError (there is no CPU fallback).

Graph           canonical lower triangle L (btc_graph_*)
BitBSR          16 x K bit tiles of L and the reference-semantics count
Plan            the hybrid engine (dense tiles + LDS-hashed wedge checks)
count_cpu       the host verify path (reference -v 1)
gen_*           seeded synthetic generators
"""
import ctypes as C
import os

import numpy as np

__all__ = [
    "BtcError", "lib", "lib_path", "Graph", "BitBSR", "Plan", "Batch", "CountOpts", "count_cpu",
    "batch_workload", "serving_workload", "Server", "autotune", "tune_lookup", "tune_table_save", "tune_table_load",
    "tune_table_clear",
    "gen_rmat", "gen_er", "gen_stencil2d", "gen_stencil3d", "gen_planted", "lite_tile_k",
    "default_tau", "version", "CANON_COMPACT", "CANON_DEGREE_ORDER", "CANON_DEFAULT", "CANON_DEVICE",
    "VARIANT_AUTO", "VARIANT_BINSEARCH", "VARIANT_DIRECT",
]

CANON_COMPACT = 1
CANON_DEGREE_ORDER = 2
CANON_DEFAULT = 3
CANON_DEVICE = 4
VARIANT_AUTO = 0
VARIANT_BINSEARCH = 1
VARIANT_DIRECT = 2
DENSE_MFMA = 16

_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(_HERE, "lib", "libbtc_amd.so")
_lib = None

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
vp = C.c_void_p


class BtcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("btc error %d: %s" % (code, msg))
        self.code = code


class GraphInfo(C.Structure):
    _fields_ = [("n", C.c_int32), ("n_input", C.c_int32), ("m", C.c_int64), ("max_row", C.c_int32),
                ("flags", C.c_uint32), ("wedges", C.c_int64), ("merge_work", C.c_int64)]


class BitBSRView(C.Structure):
    _fields_ = [("indptr", vp), ("indices", vp), ("row_indices", vp), ("tiles", vp), ("n", C.c_int32),
                ("n_row_blocks", C.c_int32), ("num_blocks", C.c_uint32), ("tile_k", C.c_int32),
                ("words_per_tile", C.c_int32)]


class CountOpts(C.Structure):
    _fields_ = [("tile_k", C.c_int32), ("tau", C.c_int32), ("chunk_work", C.c_int32),
                ("waves_per_eu", C.c_int32), ("lookup", C.c_int32), ("group_lanes", C.c_int32),
                ("unroll", C.c_int32), ("schedule", C.c_int32), ("dense_kernel", C.c_int32),
                ("planner", C.c_int32), ("core_ids", C.c_int32), ("tiny_items", C.c_int32),
                ("wide_offsets", C.c_int32)]


class PlanInfo(C.Structure):
    _fields_ = [("sparse_edges", C.c_int64), ("dense_edges", C.c_int64), ("dense_tiles", C.c_int64),
                ("wedge_checks", C.c_int64), ("work_items", C.c_int64), ("device_bytes", C.c_int64),
                ("sparse_merge_work", C.c_int64), ("stream_bytes", C.c_int64), ("core_lists", C.c_int64),
                ("core_rows", C.c_int64), ("core_ids", C.c_int64), ("tile_k", C.c_int32), ("tau", C.c_int32),
                ("tiny_items", C.c_int64)]


class PipelineInfo(C.Structure):
    _fields_ = [("n_cap", C.c_int32), ("graph_nodes", C.c_int32), ("nnz_cap", C.c_int64),
                ("device_bytes", C.c_int64), ("flags", C.c_uint32), ("n", C.c_int32), ("m", C.c_int64),
                ("wedges", C.c_int64), ("wedge_checks", C.c_int64), ("work_items", C.c_int32),
                ("max_row", C.c_int32), ("max_graphs", C.c_int32), ("reserved", C.c_int32)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_batch", C.c_int32), ("prep_threads", C.c_int32),
                ("canon_flags", C.c_int32), ("group_graphs", C.c_int32), ("captured_prep", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("graph_key", C.c_uint64), ("triangles", C.c_uint64),
                ("status", C.c_int32), ("resident_hit", C.c_int32), ("batch_size", C.c_int32),
                ("reserved", C.c_int32), ("t_submit", C.c_double), ("t_start", C.c_double),
                ("t_done", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [(k, C.c_int64) for k in ("submitted", "completed", "failed", "resident_hits", "prepared",
                                         "evictions", "launches", "launched_requests", "resident_bytes",
                                         "resident_graphs")] + [("clock", C.c_double)]


class TuneOpts(C.Structure):
    _fields_ = [("budget_ms", C.c_int32), ("reps", C.c_int32), ("tile_ks_mask", C.c_int32),
                ("reserved", C.c_int32)]


class BatchInfo(C.Structure):
    _fields_ = [("graphs", C.c_int32), ("n", C.c_int32), ("m", C.c_int64), ("plan", PlanInfo)]


_SIGS = {
    "btc_last_error": (C.c_char_p, []),
    "btc_version": (C.c_char_p, []),
    "btc_device_init": (C.c_int, [C.c_int32]),
    "btc_gen_rmat": (C.c_int64, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_uint64, vp, vp]),
    "btc_gen_er": (C.c_int64, [C.c_int32, C.c_double, C.c_uint64, vp, vp]),
    "btc_gen_stencil2d": (C.c_int64, [C.c_int32, C.c_int32, vp, vp]),
    "btc_gen_stencil3d": (C.c_int64, [C.c_int32, C.c_int32, C.c_int32, vp, vp]),
    "btc_gen_planted": (C.c_int64, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_uint64, vp, vp]),
    "btc_graph_from_coo": (C.c_int, [vp, vp, C.c_int64, C.c_int32, C.c_uint32, C.POINTER(vp)]),
    "btc_graph_from_coo_device": (C.c_int, [vp, vp, C.c_int64, C.c_int32, C.c_uint32, vp, C.POINTER(vp)]),
    "btc_graph_from_csr": (C.c_int, [vp, vp, C.c_int32, C.POINTER(vp)]),
    "btc_graph_get_info": (C.c_int, [vp, C.POINTER(GraphInfo)]),
    "btc_graph_copy_csr": (C.c_int, [vp, vp, vp]),
    "btc_graph_copy_labels": (C.c_int, [vp, vp]),
    "btc_graph_free": (None, [vp]),
    "btc_bitbsr_build": (C.c_int, [vp, C.c_int32, vp, C.POINTER(vp)]),
    "btc_bitbsr_get_view": (C.c_int, [vp, C.POINTER(BitBSRView)]),
    "btc_bitbsr_copy_to_host": (C.c_int, [vp, vp, vp, vp, vp]),
    "btc_bitbsr_free": (None, [vp]),
    "btc_bitbsr_count": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, vp, C.POINTER(C.c_uint64), vp,
                                   C.POINTER(C.c_float)]),
    "btc_bitbsr_count64": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, vp, C.POINTER(C.c_uint64), vp,
                                     C.POINTER(C.c_float)]),
    "btc_lite_tile_k": (C.c_int32, [C.c_int32, C.c_int64, C.c_int32]),
    "btc_default_tau": (C.c_int32, [C.c_int32]),
    "btc_count_opts_default": (None, [C.POINTER(CountOpts)]),
    "btc_plan_create": (C.c_int, [vp, C.POINTER(CountOpts), vp, C.POINTER(vp)]),
    "btc_plan_get_info": (C.c_int, [vp, C.POINTER(PlanInfo)]),
    "btc_plan_tune": (C.c_int, [vp, C.POINTER(CountOpts)]),
    "btc_plan_count": (C.c_int, [vp, vp, C.POINTER(C.c_uint64), vp, C.POINTER(C.c_float)]),
    "btc_plan_count64": (C.c_int, [vp, vp, C.POINTER(C.c_uint64), vp, C.POINTER(C.c_float)]),
    "btc_plan_count_noalloc": (C.c_int, [vp, vp, C.POINTER(C.c_uint64), vp]),
    "btc_plan_last_timing": (C.c_int, [vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "btc_plan_free": (None, [vp]),
    "btc_plan_capture": (C.c_int, [vp, vp]),
    "btc_plan_count_share": (C.c_int, [vp, C.c_int32, C.c_int32, vp, C.POINTER(C.c_uint64), C.POINTER(C.c_float)]),
    "btc_batch_capture": (C.c_int, [vp, vp]),
    "btc_batch_autotune": (C.c_int, [vp, C.POINTER(TuneOpts), vp, C.POINTER(CountOpts), C.POINTER(C.c_float)]),
    "btc_batch_from_coo": (C.c_int, [vp, vp, vp, vp, C.c_int32, C.c_uint32, C.POINTER(CountOpts), vp,
                                     C.POINTER(vp)]),
    "btc_count": (C.c_int, [vp, C.POINTER(CountOpts), vp, C.POINTER(C.c_uint64), vp, C.POINTER(C.c_float)]),
    "btc_count_cpu": (C.c_int, [vp, C.c_int32, C.POINTER(C.c_uint64), vp]),
    "btc_batch_create": (C.c_int, [vp, C.c_int32, C.POINTER(CountOpts), vp, C.POINTER(vp)]),
    "btc_batch_get_info": (C.c_int, [vp, C.POINTER(BatchInfo)]),
    "btc_batch_tune": (C.c_int, [vp, C.POINTER(CountOpts)]),
    "btc_batch_count": (C.c_int, [vp, vp, vp, C.POINTER(C.c_uint64), C.POINTER(C.c_float)]),
    "btc_batch_free": (None, [vp]),
    "btc_tune_opts_default": (None, [C.POINTER(TuneOpts)]),
    "btc_autotune": (C.c_int, [vp, C.POINTER(TuneOpts), vp, C.POINTER(CountOpts), C.POINTER(C.c_float)]),
    "btc_tune_lookup": (C.c_int, [vp, C.POINTER(CountOpts)]),
    "btc_tune_table_save": (C.c_int, [C.c_char_p]),
    "btc_tune_table_load": (C.c_int, [C.c_char_p]),
    "btc_tune_table_clear": (None, []),
    "btc_pipeline_create": (C.c_int, [C.c_int32, C.c_int64, C.c_uint32, C.POINTER(CountOpts), vp, C.POINTER(vp)]),
    "btc_pipeline_count": (C.c_int, [vp, vp, vp, C.c_int64, C.c_int32, C.c_int32, vp, C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_float)]),
    "btc_pipeline_create_batch": (C.c_int, [C.c_int32, C.c_int64, C.c_int32, C.c_uint32, C.POINTER(CountOpts), vp,
                                            C.POINTER(vp)]),
    "btc_pipeline_count_batch": (C.c_int, [vp, vp, vp, vp, vp, C.c_int32, vp, vp, C.POINTER(C.c_uint64),
                                           C.POINTER(C.c_float)]),
    "btc_pipeline_export_batch": (C.c_int, [vp, vp, C.POINTER(vp)]),
    "btc_pipeline_graph": (C.c_int, [vp, vp, C.POINTER(vp)]),
    "btc_pipeline_export": (C.c_int, [vp, vp, C.POINTER(vp), C.POINTER(vp)]),
    "btc_pipeline_get_info": (C.c_int, [vp, C.POINTER(PipelineInfo)]),
    "btc_pipeline_free": (None, [vp]),
    "btc_server_opts_default": (None, [C.POINTER(ServerOpts)]),
    "btc_server_create": (C.c_int, [C.POINTER(ServerOpts), C.POINTER(vp)]),
    "btc_server_submit": (C.c_int, [vp, C.c_uint64, vp, vp, C.c_int64, C.c_int32, C.POINTER(C.c_uint64)]),
    "btc_server_poll": (C.c_int, [vp, C.POINTER(Result), C.c_int32, C.c_int32, C.POINTER(C.c_int32)]),
    "btc_server_get_stats": (C.c_int, [vp, C.POINTER(ServerStats)]),
    "btc_server_free": (None, [vp]),
}


def lib():
    """Loads libbtc_amd.so once; raises BtcError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(lib_path):
            raise BtcError(-1, "libbtc_amd.so not built (run __graft_entry__.build())")
        L = C.CDLL(lib_path)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise BtcError(rc, lib().btc_last_error().decode())


def _ptr(a):
    return a.ctypes.data_as(vp) if a is not None and a.size else None


def version():
    return lib().btc_version().decode()


def lite_tile_k(n, m, mode=0):
    return int(lib().btc_lite_tile_k(n, m, mode))


def default_tau(k):
    return int(lib().btc_default_tau(k))


def _gen(fn, *args):
    cnt = fn(*args, None, None)
    if cnt < 0:
        raise BtcError(1, lib().btc_last_error().decode())
    src = np.empty(cnt, np.int32)
    dst = np.empty(cnt, np.int32)
    fn(*args, _ptr(src), _ptr(dst))
    return src, dst


def gen_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
    return _gen(lib().btc_gen_rmat, scale, edge_factor, a, b, c, seed)


def gen_er(n, avg_degree, seed=1):
    return _gen(lib().btc_gen_er, n, avg_degree, seed)


def gen_stencil2d(nx, ny):
    return _gen(lib().btc_gen_stencil2d, nx, ny)


def gen_stencil3d(nx, ny, nz):
    return _gen(lib().btc_gen_stencil3d, nx, ny, nz)


def gen_planted(n, community, p_in, noise_degree, seed=1):
    return _gen(lib().btc_gen_planted, n, community, p_in, noise_degree, seed)


class Graph:
    """Canonical strictly-lower L of an undirected graph (host + HBM)."""

    def __init__(self, handle):
        self._h = vp(handle)

    @classmethod
    def from_coo(cls, src, dst, n=0, flags=CANON_DEFAULT):
        src = np.ascontiguousarray(src, np.int32)
        dst = np.ascontiguousarray(dst, np.int32)
        if src.shape != dst.shape:
            raise ValueError("src and dst differ in length")
        h = vp()
        _check(lib().btc_graph_from_coo(_ptr(src), _ptr(dst), src.size, n, flags, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_coo_device(cls, src_ptr, dst_ptr, nnz, n=0, flags=CANON_DEFAULT, stream=None):
        """From int32 pairs already in HBM (device addresses, e.g. a torch
        tensor's data_ptr()), canonicalised on the GPU."""
        h = vp()
        _check(lib().btc_graph_from_coo_device(vp(src_ptr), vp(dst_ptr), nnz, n, flags, stream, C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_csr(cls, rowptr, col):
        rowptr = np.ascontiguousarray(rowptr, np.int64)
        col = np.ascontiguousarray(col, np.int32)
        h = vp()
        _check(lib().btc_graph_from_csr(_ptr(rowptr), _ptr(col), rowptr.size - 1, C.byref(h)))
        return cls(h.value)

    @property
    def info(self):
        gi = GraphInfo()
        _check(lib().btc_graph_get_info(self._h, C.byref(gi)))
        return {k: getattr(gi, k) for k, _ in GraphInfo._fields_}

    @property
    def n(self):
        return self.info["n"]

    @property
    def m(self):
        return self.info["m"]

    def csr(self):
        i = self.info
        rp = np.empty(i["n"] + 1, np.int64)
        col = np.empty(i["m"], np.int32)
        _check(lib().btc_graph_copy_csr(self._h, _ptr(rp), _ptr(col)))
        return rp, col

    def labels(self):
        lab = np.empty(self.info["n"], np.int32)
        _check(lib().btc_graph_copy_labels(self._h, _ptr(lab)))
        return lab

    def close(self):
        if self._h is not None and self._h.value:
            lib().btc_graph_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def _count_out(n, per_vertex, wide=False):
    total = C.c_uint64(0)
    ms = C.c_float(0)
    pv = np.zeros(n, np.int64 if wide else np.int32) if per_vertex else None
    return total, ms, pv


class BitBSR:
    """16 x K Bit-BSR of a Graph, resident in HBM."""

    def __init__(self, graph, tile_k, stream=None):
        self.graph = graph
        h = vp()
        _check(lib().btc_bitbsr_build(graph._h, tile_k, stream, C.byref(h)))
        self._h = h

    @property
    def view(self):
        v = BitBSRView()
        _check(lib().btc_bitbsr_get_view(self._h, C.byref(v)))
        return {k: getattr(v, k) for k, _ in BitBSRView._fields_}

    def to_host(self):
        v = self.view
        indptr = np.empty(v["n_row_blocks"] + 1, np.uint32)
        nb = v["num_blocks"]
        indices = np.empty(nb, np.uint32)
        rows = np.empty(nb, np.uint32)
        tiles = np.empty(nb * v["words_per_tile"], np.uint64)
        _check(lib().btc_bitbsr_copy_to_host(self._h, _ptr(indptr), _ptr(indices), _ptr(rows), _ptr(tiles)))
        return indptr, indices, rows, tiles

    def count(self, variant=VARIANT_AUTO, threshold_blocks=0, tau=-1, per_vertex=False, stream=None, mfma=False,
              wide=False):
        """Returns (total, per_vertex or None, kernel_ms); mfma=True counts
        dense tiles with the i8 MFMA form; wide=True returns int64
        per-vertex counts (btc_bitbsr_count64)."""
        total, ms, pv = _count_out(self.view["n"], per_vertex, wide)
        if mfma:
            variant |= DENSE_MFMA
        fn = lib().btc_bitbsr_count64 if wide else lib().btc_bitbsr_count
        _check(fn(self._h, variant, threshold_blocks, tau, stream, C.byref(total), _ptr(pv), C.byref(ms)))
        return total.value, pv, ms.value

    def close(self):
        if self._h is not None and self._h.value:
            lib().btc_bitbsr_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_opts(tile_k=None, tau=None, chunk_work=None, waves_per_eu=None, lookup=None, group_lanes=None,
              unroll=None, schedule=None, dense_kernel=None, planner=None, core_ids=None, tiny_items=None,
              wide_offsets=None):
    """btc_count_opts with the library defaults, overridden where given."""
    o = CountOpts()
    lib().btc_count_opts_default(C.byref(o))
    for k, v in dict(tile_k=tile_k, tau=tau, chunk_work=chunk_work, waves_per_eu=waves_per_eu, lookup=lookup,
                     group_lanes=group_lanes, unroll=unroll, schedule=schedule,
                     dense_kernel=dense_kernel, planner=planner, core_ids=core_ids,
                     tiny_items=tiny_items, wide_offsets=wide_offsets).items():
        if v is not None:
            setattr(o, k, v)
    return o


class Plan:
    """Hybrid engine plan for one Graph (device-resident work lists)."""

    def __init__(self, graph, stream=None, **opts):
        self.graph = graph
        self.opts = make_opts(**opts)
        h = vp()
        _check(lib().btc_plan_create(graph._h, C.byref(self.opts), stream, C.byref(h)))
        self._h = h

    @property
    def info(self):
        pi = PlanInfo()
        _check(lib().btc_plan_get_info(self._h, C.byref(pi)))
        return {k: getattr(pi, k) for k, _ in PlanInfo._fields_}

    def tune(self, **opts):
        """Changes launch-only options (waves_per_eu, lookup, group_lanes, unroll, schedule, dense_kernel)."""
        o = make_opts(**opts)
        _check(lib().btc_plan_tune(self._h, C.byref(o)))

    def count(self, per_vertex=False, stream=None, wide=False):
        """Returns (total, per_vertex or None, kernel_ms); wide=True returns
        int64 per-vertex counts (btc_plan_count64)."""
        total, ms, pv = _count_out(self.graph.n, per_vertex, wide)
        fn = lib().btc_plan_count64 if wide else lib().btc_plan_count
        _check(fn(self._h, stream, C.byref(total), _ptr(pv), C.byref(ms)))
        return total.value, pv, ms.value

    def count_noalloc(self, per_vertex=False, stream=None):
        """(total, per_vertex or None) from btc_plan_count_noalloc: the count
        in the library's static device workspace, allocating nothing."""
        total = C.c_uint64(0)
        pv = np.zeros(self.graph.n, np.int32) if per_vertex else None
        _check(lib().btc_plan_count_noalloc(self._h, stream, C.byref(total), _ptr(pv)))
        return total.value, pv

    def count_share(self, shares, share, stream=None):
        """(total, kernel_ms) of share `share` of a `shares`-way split."""
        total, ms = C.c_uint64(0), C.c_float(0)
        _check(lib().btc_plan_count_share(self._h, shares, share, stream, C.byref(total), C.byref(ms)))
        return total.value, ms.value

    def capture(self, stream=None):
        """Captures the totals-only count into a HIP graph that later
        count() calls without per_vertex replay."""
        _check(lib().btc_plan_capture(self._h, stream))

    def last_timing(self):
        """(dense-tile kernel ms, wedge kernel ms) of the last count (-1 after a replay)."""
        dm, wm = C.c_float(0), C.c_float(0)
        _check(lib().btc_plan_last_timing(self._h, C.byref(dm), C.byref(wm)))
        return dm.value, wm.value

    def close(self):
        if self._h is not None and self._h.value:
            lib().btc_plan_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pipeline:
    """COO pairs -> triangle total as one captured HIP graph, for inputs of
    up to n_cap ids and nnz_cap pairs (btc_pipeline)."""

    def __init__(self, n_cap, nnz_cap, flags=CANON_DEFAULT, stream=None, max_graphs=1, **opts):
        self.opts = make_opts(**opts)
        h = vp()
        _check(lib().btc_pipeline_create_batch(n_cap, nnz_cap, max_graphs, flags, C.byref(self.opts), stream,
                                               C.byref(h)))
        self._h = h

    def count_batch(self, graphs, stream=None):
        """(total, per-graph uint64 array, kernel_ms) of [(src, dst, n), ...]
        in one run (btc_pipeline_count_batch)."""
        k = len(graphs)
        off = np.zeros(k + 1, np.int64)
        off[1:] = np.cumsum([len(g[0]) for g in graphs])
        src = np.ascontiguousarray(np.concatenate([np.asarray(g[0], np.int32) for g in graphs]), np.int32)
        dst = np.ascontiguousarray(np.concatenate([np.asarray(g[1], np.int32) for g in graphs]), np.int32)
        nv = np.ascontiguousarray([g[2] for g in graphs], np.int32)
        per = np.zeros(k, np.uint64)
        total, ms = C.c_uint64(0), C.c_float(0)
        _check(lib().btc_pipeline_count_batch(self._h, _ptr(src), _ptr(dst), _ptr(off), _ptr(nv), k, stream,
                                              _ptr(per), C.byref(total), C.byref(ms)))
        return total.value, per, ms.value

    def export_batch(self, stream=None):
        """The last run of several graphs as a Batch."""
        h = vp()
        _check(lib().btc_pipeline_export_batch(self._h, stream, C.byref(h)))
        b = Batch.__new__(Batch)
        b.n_graphs = self.info["max_graphs"]
        b.opts = self.opts
        b._h = h
        b.n_graphs = b.info["graphs"]
        return b

    def count(self, src, dst, n=0, stream=None):
        """(total, kernel_ms) of the graph of host pairs (src[e], dst[e])."""
        src = np.ascontiguousarray(src, np.int32)
        dst = np.ascontiguousarray(dst, np.int32)
        if src.shape != dst.shape:
            raise ValueError("src and dst differ in length")
        total, ms = C.c_uint64(0), C.c_float(0)
        _check(lib().btc_pipeline_count(self._h, _ptr(src), _ptr(dst), src.size, n, 0, stream, C.byref(total),
                                        C.byref(ms)))
        return total.value, ms.value

    def count_device(self, src_ptr, dst_ptr, nnz, n=0, stream=None):
        """(total, kernel_ms) of int32 pairs already in HBM (device addresses)."""
        total, ms = C.c_uint64(0), C.c_float(0)
        _check(lib().btc_pipeline_count(self._h, vp(src_ptr), vp(dst_ptr), nnz, n, 1, stream, C.byref(total),
                                        C.byref(ms)))
        return total.value, ms.value

    def graph(self, stream=None):
        """The last run's canonical graph (a Graph in HBM)."""
        h = vp()
        _check(lib().btc_pipeline_graph(self._h, stream, C.byref(h)))
        return Graph(h.value)

    def export(self, stream=None):
        """(Graph, Plan) of the last run, in buffers of their own."""
        g, q = vp(), vp()
        _check(lib().btc_pipeline_export(self._h, stream, C.byref(g), C.byref(q)))
        graph = Graph(g.value)
        plan = Plan.__new__(Plan)
        plan.graph = graph
        plan.opts = self.opts
        plan._h = q
        return graph, plan

    @property
    def info(self):
        pi = PipelineInfo()
        _check(lib().btc_pipeline_get_info(self._h, C.byref(pi)))
        return {k: getattr(pi, k) for k, _ in PipelineInfo._fields_}

    def close(self):
        if self._h is not None and self._h.value:
            lib().btc_pipeline_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def count_cpu(graph, threads=0, per_vertex=False):
    total, _, pv = _count_out(graph.n, per_vertex)
    _check(lib().btc_count_cpu(graph._h, threads, C.byref(total), _ptr(pv)))
    return total.value, pv


class Batch:
    """Many graphs counted by one persistent launch (btc_batch_*)."""

    def __init__(self, graphs, stream=None, **opts):
        self.n_graphs = len(graphs)
        self.opts = make_opts(**opts)
        arr = (vp * self.n_graphs)(*[g._h.value for g in graphs])
        h = vp()
        _check(lib().btc_batch_create(arr, self.n_graphs, C.byref(self.opts), stream, C.byref(h)))
        self._h = h

    @classmethod
    def from_coo(cls, graphs, flags=CANON_DEFAULT, stream=None, **opts):
        """A batch built on the GPU straight from [(src, dst, n), ...]
        (btc_batch_from_coo)."""
        self = cls.__new__(cls)
        self.n_graphs = len(graphs)
        self.opts = make_opts(**opts)
        off = np.zeros(self.n_graphs + 1, np.int64)
        off[1:] = np.cumsum([len(g[0]) for g in graphs])
        src = np.ascontiguousarray(np.concatenate([np.asarray(g[0], np.int32) for g in graphs]), np.int32)
        dst = np.ascontiguousarray(np.concatenate([np.asarray(g[1], np.int32) for g in graphs]), np.int32)
        nv = np.ascontiguousarray([g[2] for g in graphs], np.int32)
        h = vp()
        _check(lib().btc_batch_from_coo(_ptr(src), _ptr(dst), _ptr(off), _ptr(nv), self.n_graphs, flags,
                                        C.byref(self.opts), stream, C.byref(h)))
        self._h = h
        return self

    @property
    def info(self):
        bi = BatchInfo()
        _check(lib().btc_batch_get_info(self._h, C.byref(bi)))
        d = {k: getattr(bi, k) for k, _ in BatchInfo._fields_ if k != "plan"}
        d["plan"] = {k: getattr(bi.plan, k) for k, _ in PlanInfo._fields_}
        return d

    def tune(self, **opts):
        o = make_opts(**opts)
        _check(lib().btc_batch_tune(self._h, C.byref(o)))

    def capture(self, stream=None):
        """Captures the batch count into a HIP graph that later count() calls replay."""
        _check(lib().btc_batch_capture(self._h, stream))

    def autotune(self, budget_ms=None, reps=None, tile_ks_mask=None, stream=None):
        """Tunes the batch's options on the device and re-plans it; returns
        (best options dict, ms)."""
        t = TuneOpts()
        lib().btc_tune_opts_default(C.byref(t))
        for k, v in dict(budget_ms=budget_ms, reps=reps, tile_ks_mask=tile_ks_mask).items():
            if v is not None:
                setattr(t, k, v)
        best = CountOpts()
        ms = C.c_float(0)
        _check(lib().btc_batch_autotune(self._h, C.byref(t), stream, C.byref(best), C.byref(ms)))
        return _opts_dict(best), ms.value

    def count(self, stream=None):
        """Returns (total, per-graph uint64 array, kernel_ms)."""
        per = np.zeros(self.n_graphs, np.uint64)
        total = C.c_uint64(0)
        ms = C.c_float(0)
        _check(lib().btc_batch_count(self._h, stream, _ptr(per), C.byref(total), C.byref(ms)))
        return total.value, per, ms.value

    def close(self):
        if self._h is not None and self._h.value:
            lib().btc_batch_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def batch_workload(count=4096, seed=1, n_lo=256, n_hi=16384):
    """The batched workload: `count` graphs with n log-uniform in
    [n_lo, n_hi]; 40 % Erdos-Renyi (average degree uniform in [8, 32]), 40 %
    R-MAT (scale ceil(log2 n), edge factor uniform in {8..16}, a, b, c =
    0.57, 0.19, 0.19; pairs with an id >= n are dropped), 20 % 9-point
    stencils (a x b, a = floor(sqrt(n)), b = n // a).  The spec is drawn with
    numpy's PCG64 from `seed`; graph k's generator seed is drawn with it.
    Returns a list of (src, dst, n, kind)."""
    import math
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        r = rng.random()
        gseed = int(rng.integers(1, 2 ** 62))
        if r < 0.4:
            src, dst = gen_er(n, float(rng.uniform(8, 32)), gseed)
            kind = "er"
        elif r < 0.8:
            scale = max(1, math.ceil(math.log2(n)))
            ef = int(rng.integers(8, 17))
            src, dst = gen_rmat(scale, ef, 0.57, 0.19, 0.19, gseed)
            keep = (src < n) & (dst < n)
            src, dst = src[keep], dst[keep]
            kind = "rmat"
        else:
            a = int(math.isqrt(n))
            b = n // a
            src, dst = gen_stencil2d(a, b)
            n = a * b
            kind = "stencil2d"
        out.append((src, dst, n, kind))
    return out


def serving_workload(requests, seed=1, key0=0, repeat=0.3, median_n=1 << 13, sigma=1.0, n_cap=1 << 20,
                     edge_factor=16):
    """The serving request stream: `requests` exact-count queries.  With
    probability `repeat` a query repeats a graph key drawn uniformly from
    the keys submitted before it (no edges sent); otherwise it brings a new
    R-MAT graph (edge factor 16, a, b, c = 0.57, 0.19, 0.19, scale ceil(log2
    n), pairs with an id >= n dropped) with n log-normal (median 2^13, sigma
    1 in ln n, at least 64, capped at 2^20).  Draws use numpy's PCG64 from
    `seed`.  Returns a list of (key, (src, dst, n) or None); new keys are
    key0, key0 + 1, ..."""
    import math
    rng = np.random.default_rng(seed)
    reqs, keys = [], []
    for _ in range(requests):
        if keys and rng.random() < repeat:
            reqs.append((int(rng.choice(keys)), None))
            continue
        n = int(min(n_cap, max(64, round(math.exp(rng.normal(math.log(median_n), sigma))))))
        sc = max(1, math.ceil(math.log2(n)))
        s, d = gen_rmat(sc, edge_factor, 0.57, 0.19, 0.19, seed=int(rng.integers(1, 2 ** 62)))
        keep = (s < n) & (d < n)
        key = key0 + len(keys)
        keys.append(key)
        reqs.append((key, (s[keep], d[keep], n)))
    return reqs


class Server:
    """Request loop with continuous batching (btc_server_*)."""

    def __init__(self, hbm_budget_bytes=None, max_batch=None, prep_threads=None, canon_flags=None,
                 group_graphs=None, captured_prep=None):
        o = ServerOpts()
        lib().btc_server_opts_default(C.byref(o))
        for k, v in dict(hbm_budget_bytes=hbm_budget_bytes, max_batch=max_batch, prep_threads=prep_threads,
                         canon_flags=canon_flags, group_graphs=group_graphs, captured_prep=captured_prep).items():
            if v is not None:
                setattr(o, k, v)
        h = vp()
        _check(lib().btc_server_create(C.byref(o), C.byref(h)))
        self._h = h
        self._buf = (Result * 4096)()

    def submit(self, key, src=None, dst=None, n=0):
        """Queues a count of graph `key`; the edges are copied (omit them for
        a key submitted before).  Returns the request id."""
        rid = C.c_uint64(0)
        if src is None:
            _check(lib().btc_server_submit(self._h, key, None, None, 0, n, C.byref(rid)))
        else:
            src = np.ascontiguousarray(src, np.int32)
            dst = np.ascontiguousarray(dst, np.int32)
            _check(lib().btc_server_submit(self._h, key, _ptr(src), _ptr(dst), src.size, n, C.byref(rid)))
        return rid.value

    def poll(self, wait_ms=0):
        got = C.c_int32(0)
        _check(lib().btc_server_poll(self._h, self._buf, len(self._buf), wait_ms, C.byref(got)))
        return [{k: getattr(self._buf[i], k) for k, _ in Result._fields_ if k != "reserved"}
                for i in range(got.value)]

    def stats(self):
        st = ServerStats()
        _check(lib().btc_server_get_stats(self._h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self._h is not None and self._h.value:
            lib().btc_server_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _opts_dict(o):
    return {k: getattr(o, k) for k, _ in CountOpts._fields_}


def autotune(graph, budget_ms=None, reps=None, tile_ks_mask=None, stream=None):
    """Times candidate options on the device; returns (best options dict, ms)."""
    t = TuneOpts()
    lib().btc_tune_opts_default(C.byref(t))
    for k, v in dict(budget_ms=budget_ms, reps=reps, tile_ks_mask=tile_ks_mask).items():
        if v is not None:
            setattr(t, k, v)
    best = CountOpts()
    ms = C.c_float(0)
    _check(lib().btc_autotune(graph._h, C.byref(t), stream, C.byref(best), C.byref(ms)))
    return _opts_dict(best), ms.value


def tune_lookup(graph):
    """Tuned options for the graph's feature class, or None."""
    o = CountOpts()
    rc = lib().btc_tune_lookup(graph._h, C.byref(o))
    return _opts_dict(o) if rc == 0 else None


def tune_table_save(path):
    _check(lib().btc_tune_table_save(str(path).encode()))


def tune_table_load(path):
    _check(lib().btc_tune_table_load(str(path).encode()))


def tune_table_clear():
    lib().btc_tune_table_clear()
