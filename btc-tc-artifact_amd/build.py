"""Compiles libbtc_amd.so (HIP kernels for gfx950 + host C++) in-tree.

Usage: python build.py [--jobs N] [--verbose]
The library lands in btc_amd/lib/libbtc_amd.so next to the Python binding.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "btc_amd", "lib")
LIB = os.path.join(LIB_DIR, "libbtc_amd.so")
BIN_DIR = os.path.join(HERE, "bin")
APPS = os.path.join(HERE, "apps")

ARCH = os.environ.get("BTC_OFFLOAD_ARCH", "gfx950")
COMMON = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-Wall", "-Wno-unused-result",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC]
# packed-fp32 instructions stay off for every device object
HIP_ONLY = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-munsafe-fp-atomics"]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith((".cpp", ".hip")):
            out.append(os.path.join(CSRC, f))
    return out


def compile_one(src, verbose):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    # hipcc compiles .cpp as HIP too; give every object the same target
    cmd = [hipcc()] + COMMON + HIP_ONLY
    if src.endswith(".hip"):
        cmd += ["-x", "hip"]
    cmd += ["-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # hipcc's host pass does not know the packed-fp32 feature; that note is harmless
    msg = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (src, msg))
    if msg.strip() and verbose:
        print(msg)
    return obj


def build(jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: compile_one(s, verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fopenmp", "-pthread", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed\n" + r.stdout)
    os.replace(tmp, LIB)
    build_apps(verbose)
    return LIB


def build_app(src, out, verbose=False):
    """Links a host C++ program against libbtc_amd.so (rpath to btc_amd/lib)."""
    cxx = os.environ.get("CXX", "g++")
    cmd = [cxx, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", HERE, src, "-o", out,
           "-L", LIB_DIR, "-lbtc_amd", "-Wl,-rpath," + LIB_DIR]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("app build failed: %s\n%s" % (src, r.stdout))
    return out


def build_apps(verbose=False):
    os.makedirs(BIN_DIR, exist_ok=True)
    for f in sorted(os.listdir(APPS)):
        if f.endswith(".cpp"):
            build_app(os.path.join(APPS, f), os.path.join(BIN_DIR, f[:-4]), verbose)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
