"""ctypes binding of the CPU oracle (liborc.so) and of the reference's own CPU
verifier compiled from its header (_ref/libreftc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  See oracle.c for the
reference file:line each function restates and how each is pinned.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")
REF_LIB = os.path.join(HERE, "_ref", "libreftc.so")
vp = C.c_void_p
_orc = None
_ref = None


def build(with_ref=None):
    """Compiles liborc.so (and _ref/libreftc.so when the reference tree exists)."""
    targets = ["all"]
    if with_ref is None:
        with_ref = os.path.isdir("/root/reference/btc")
    if with_ref:
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def _lib():
    global _orc
    if _orc is None:
        if not os.path.exists(LIB):
            build(with_ref=False)
        L = C.CDLL(LIB)
        L.orc_canon.restype = C.c_int64
        L.orc_canon.argtypes = [vp, vp, C.c_int64, C.c_int32, C.c_uint32, C.POINTER(vp), C.POINTER(vp),
                                C.POINTER(vp), C.POINTER(C.c_int32)]
        L.orc_free.argtypes = [vp]
        L.orc_tc_merge.restype = C.c_uint64
        L.orc_tc_merge.argtypes = [vp, vp, C.c_int32, C.c_int32, vp]
        L.orc_tc_brute.restype = C.c_uint64
        L.orc_tc_brute.argtypes = [vp, vp, C.c_int32, vp]
        L.orc_bitbsr_build.restype = C.c_int64
        L.orc_bitbsr_build.argtypes = [vp, vp, C.c_int32, C.c_int32, vp, vp, vp, vp]
        L.orc_bitbsr_count.restype = C.c_uint64
        L.orc_bitbsr_count.argtypes = [C.c_int32, C.c_int32, vp, vp, vp, vp, C.c_uint32, C.c_int32, C.c_int32,
                                       vp, C.POINTER(C.c_int64)]
        L.orc_lite_tile_k.restype = C.c_int32
        L.orc_lite_tile_k.argtypes = [C.c_int32, C.c_int64, C.c_int32]
        L.orc_variant.restype = C.c_int32
        L.orc_variant.argtypes = [C.c_int32, C.c_uint32, C.c_int32]
        _orc = L
    return _orc


def _p(a):
    return a.ctypes.data_as(vp) if a is not None and a.size else None


def canon(src, dst, n=0, flags=3):
    """Canonical L of (src, dst): (rowptr int64, col int32, labels int32)."""
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    rp, cl, lb = vp(), vp(), vp()
    nc = C.c_int32(0)
    m = _lib().orc_canon(_p(src), _p(dst), src.size, n, flags, C.byref(rp), C.byref(cl), C.byref(lb),
                         C.byref(nc))
    if m < 0:
        raise ValueError("oracle canon: id out of range")
    n_out = nc.value
    rowptr = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_int64)), (n_out + 1,)).copy()
    col = np.ctypeslib.as_array(C.cast(cl, C.POINTER(C.c_int32)), (max(m, 1),))[:m].copy()
    labels = np.ctypeslib.as_array(C.cast(lb, C.POINTER(C.c_int32)), (max(n_out, 1),))[:n_out].copy()
    for ptr in (rp, cl, lb):
        _lib().orc_free(ptr)
    return rowptr, col, labels


def tc_merge(rowptr, col, threads=1, per_row=False):
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    n = rowptr.size - 1
    pr = np.zeros(n, np.int64) if per_row else None
    t = _lib().orc_tc_merge(_p(rowptr), _p(col), n, threads, _p(pr))
    return (t, pr) if per_row else t


def tc_brute(rowptr, col, per_row=False):
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    n = rowptr.size - 1
    if n > 4096:
        raise ValueError("brute force is limited to n <= 4096")
    pr = np.zeros(n, np.int64) if per_row else None
    t = _lib().orc_tc_brute(_p(rowptr), _p(col), n, _p(pr))
    return (t, pr) if per_row else t


def bitbsr_build(rowptr, col, K):
    rowptr = np.ascontiguousarray(rowptr, np.int64)
    col = np.ascontiguousarray(col, np.int32)
    n = rowptr.size - 1
    nb = _lib().orc_bitbsr_build(_p(rowptr), _p(col), n, K, None, None, None, None)
    if nb < 0:
        raise ValueError("bad tile width")
    n_rb = (n + 15) // 16
    indptr = np.zeros(n_rb + 1, np.uint32)
    indices = np.zeros(max(nb, 1), np.uint32)
    rows = np.zeros(max(nb, 1), np.uint32)
    tiles = np.zeros(max(nb * (K // 4), 1), np.uint64)
    _lib().orc_bitbsr_build(_p(rowptr), _p(col), n, K, _p(indptr), _p(indices), _p(rows), _p(tiles))
    return indptr, indices[:nb], rows[:nb], tiles[:nb * (K // 4)]


def bitbsr_count(K, n, indptr, indices, rows, tiles, tau=None, variant=2, per_row=False):
    """Returns (total, per_row[n] or None, dense_tiles)."""
    if tau is None:
        tau = {128: 512, 64: 128, 32: 64}[K]
    n_rb = (n + 15) // 16
    pr = np.zeros(n_rb * 16, np.int64) if per_row else None
    nd = C.c_int64(0)
    indptr, indices, rows, tiles = (np.ascontiguousarray(a) for a in (indptr, indices, rows, tiles))
    t = _lib().orc_bitbsr_count(K, n, _p(indptr), _p(indices), _p(rows), _p(tiles), indices.size, tau, variant,
                                _p(pr), C.byref(nd))
    return t, (pr[:n] if per_row else None), nd.value


def lite_tile_k(n, m, mode=0):
    return _lib().orc_lite_tile_k(n, m, mode)


def variant(K, nblk, threshold=0):
    return _lib().orc_variant(K, nblk, threshold)


def ref_available():
    return os.path.exists(REF_LIB)


def ref_cpu_tc(rowptr, col, threads=1):
    """The reference's cpu_tc_intersection_lower_triangular on a CSR (int32 offsets)."""
    global _ref
    if _ref is None:
        if not ref_available():
            return None
        _ref = C.CDLL(REF_LIB)
        _ref.ref_cpu_tc_lower.restype = C.c_uint64
        _ref.ref_cpu_tc_lower.argtypes = [vp, vp, C.c_int32, C.c_int32]
    rp32 = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    return _ref.ref_cpu_tc_lower(_p(rp32), _p(col), rp32.size - 1, threads)
