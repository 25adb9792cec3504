"""Builds libmosaic_hip.so (gfx950) in-tree from csrc/*.hip, then the C++
host API (host/mosaic.cpp) with its test program (test_mosaic).

Flags: -ffp-contract=off keeps the reference's unfused f32 arithmetic;
packed FP32 instructions are disabled (-packed-fp32-ops target feature).
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmosaic_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MH_OFFLOAD_ARCH", "gfx950")

HIPFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *HIPFLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_host(verbose)
    return LIB


HOST = os.path.join(HERE, "host")
HOST_TESTS = os.path.join(HERE, "test_mosaic")


def build_host(verbose=False):
    """The C++ host API and its tests, linked against libmosaic_hip.so."""
    include = os.path.join(os.path.dirname(HERE), "include")
    cmd = [os.environ.get("CXX", "g++"), "-std=c++17", "-O2", "-Wall", "-Wextra", f"-I{include}",
           os.path.join(HOST, "mosaic.cpp"), os.path.join(HOST, "test_mosaic.cpp"),
           f"-L{HERE}", "-lmosaic_hip", "-Wl,-rpath,$ORIGIN", "-o", HOST_TESTS + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(HOST_TESTS + ".tmp", HOST_TESTS)
    if verbose:
        print(f"built {HOST_TESTS}")


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
