"""Host AddressSanitizer + UndefinedBehaviorSanitizer build of the CPU-side
code: the library's generators, host canonicalisation, CPU count and graph
handles (btc-tc-artifact_amd/csrc/{gen,canon,runtime}.cpp, built with g++
against the HIP headers, no device code), the readers of btc/btc.h and the
oracle (oracle/oracle.c), driven by tests/cpp/sanitize_host.cpp.  The
driver checks every result against the oracle and feeds malformed inputs;
any sanitizer report fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "btc-tc-artifact_amd")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1",
       "-fopenmp"]
HIP_INC = "/opt/rocm/include"
HIP_LIB = "/opt/rocm/lib"


@pytest.mark.skipif(not shutil.which("g++") or not os.path.isdir(HIP_INC), reason="needs g++ and the HIP headers")
def test_host_code_under_sanitizers(tmp_path):
    inc = ["-I", os.path.join(ROOT, "include"), "-I", PKG, "-I", os.path.join(PKG, "csrc"),
           "-I", os.path.join(ROOT, "oracle"), "-I", HIP_INC, "-D__HIP_PLATFORM_AMD__"]
    objs = []
    jobs = [(["gcc", "-std=c11"], os.path.join(ROOT, "oracle", "oracle.c"))]
    for f in ("gen.cpp", "canon.cpp", "runtime.cpp"):
        jobs.append((["g++", "-std=c++17"], os.path.join(PKG, "csrc", f)))
    jobs.append((["g++", "-std=c++17"], os.path.join(ROOT, "tests", "cpp", "sanitize_host.cpp")))
    procs = []
    for cc, src in jobs:
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        objs.append(obj)
        procs.append(subprocess.Popen(cc + SAN + inc + ["-c", src, "-o", obj], stdout=subprocess.PIPE,
                                      stderr=subprocess.STDOUT, text=True))
    for p, (_, src) in zip(procs, jobs):
        out = p.communicate()[0]
        assert p.returncode == 0, "compile of %s failed:\n%s" % (src, out)
    exe = str(tmp_path / "sanitize_host")
    r = subprocess.run(["g++"] + SAN + objs + ["-o", exe, "-L", HIP_LIB, "-lamdhip64", "-Wl,-rpath," + HIP_LIB],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    scratch = tmp_path / "files"
    scratch.mkdir()
    r = subprocess.run([exe, str(scratch)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, env=env,
                       timeout=600)
    assert r.returncode == 0 and "sanitize ok" in r.stdout, r.stdout[-6000:]
