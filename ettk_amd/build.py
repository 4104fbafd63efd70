"""Build libettk_align.so (hipcc, gfx950) in-tree, and the oracle/checker libraries under oracle/.

Usage: python -m ettk_amd.build [--no-oracle]
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libettk_align.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 in any kernel
HIP_FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# align.cpp runs the Powell refinements as C++20 coroutines
STD_OVERRIDE = {"align.cpp": "-std=c++20"}
SOURCES = ["fft.hip", "ccc.hip", "so3.hip", "refine.hip", "host.cpp", "align.cpp", "refine.cpp", "avg.hip", "avg.cpp",
           "noalloc.hip", "extract.hip", "classify.cpp", "tune.cpp", "cand.hip"]


def build_lib(verbose: bool = False) -> Path:
    """Compiles every source (several hipcc processes at once) and links libettk_align.so in-tree."""
    from concurrent.futures import ThreadPoolExecutor

    def compile_one(src):
        obj = CSRC / ("_" + src + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        flags = [STD_OVERRIDE.get(src, x) if x == "-std=c++17" else x for x in HIP_FLAGS]
        cmd = [HIPCC, *flags, *lang, "-c", str(CSRC / src), "-o", str(obj)]
        return src, obj, subprocess.run(cmd, capture_output=True, text=True)

    try:
        jobs = max(1, min(8, len(os.sched_getaffinity(0))))
    except AttributeError:
        jobs = max(1, min(8, os.cpu_count() or 1))
    objs = []
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(compile_one, SOURCES))
    for src, obj, r in results:
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        if verbose:
            sys.stderr.write(r.stderr)
        objs.append(str(obj))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *objs, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libettk_align.so failed")
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def build_cpp_tests() -> Path:
    """The C++ drop-in's test driver (tests/cpp/test_metric), linked against libettk_align.so."""
    src = ROOT / "tests" / "cpp" / "test_metric.cpp"
    exe = src.with_suffix("")
    cmd = ["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", str(ROOT / "include"), "-I", "/opt/rocm/include",
           str(src), "-o", str(exe), "-L", str(PKG), "-lettk_align", "-L", "/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/../../ettk_amd", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build of tests/cpp/test_metric failed")
    return exe


def build_io_test() -> Path:
    """The driver of the on-disk formats (tests/cpp/test_io): host code only."""
    src = ROOT / "tests" / "cpp" / "test_io.cpp"
    exe = src.with_suffix("")
    cmd = ["g++", "-O2", "-std=c++17", "-I", str(ROOT / "include"), str(src), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build of tests/cpp/test_io failed")
    return exe


def build_oracle() -> None:
    """The FP64 CPU oracle (test infrastructure) and, where the reference tree exists, the soft-1.0 checker."""
    subprocess.run(["make", "-C", str(ROOT / "oracle"), "-s", "liboracle.so"], check=True)
    if Path("/root/reference/external/soft-1.0").is_dir():
        subprocess.run(["make", "-C", str(ROOT / "oracle"), "-s", "ref"], check=True)


if __name__ == "__main__":
    build_lib(verbose="-v" in sys.argv)
    build_cpp_tests()
    build_io_test()
    if "--no-oracle" not in sys.argv:
        build_oracle()
    print(LIB)
