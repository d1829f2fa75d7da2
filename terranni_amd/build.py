"""Build the engine library (gfx950) and the CPU oracle, in-tree.

    python -m terranni_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "_lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["tnni_kernels.hip", "tnni_geometry.hip", "tnni_engine.hip"]
# -ffp-contract=off: every FP64 de-scale step rounds as written, so the
# device and the CPU oracle agree bit for bit.  The packed-fp32 feature is
# switched off for all device code.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in r.stdout.splitlines() if "packed-fp32-ops" not in line)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def build_lib(verbose: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libterranni_hip.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    _run([HIPCC, *HIP_FLAGS, "-shared", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
          *srcs, "-o", tmp, "-lpthread"], verbose)
    os.replace(tmp, out)
    return out


CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra"]


def build_nni_api(verbose: bool = False) -> str:
    """libterranni_nni.so: the nni.h API over the engine's C ABI (host C++),
    and the small driver program the API test runs."""
    out = os.path.join(LIB_DIR, "libterranni_nni.so")
    tmp = out + ".tmp"
    inc = ["-I", os.path.join(ROOT, "include")]
    _run([CXX, *CXX_FLAGS, "-shared", *inc, os.path.join(CSRC, "nni_shim.cpp"), "-o", tmp,
          "-L", LIB_DIR, "-lterranni_hip", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, out)
    tdir = os.path.join(ROOT, "tests", "cpp", "_build")
    os.makedirs(tdir, exist_ok=True)
    exe = os.path.join(tdir, "nni_api_test")
    _run([CXX, *CXX_FLAGS, *inc, os.path.join(ROOT, "tests", "cpp", "nni_api_test.cpp"), "-o",
          exe + ".tmp", "-L", LIB_DIR, "-lterranni_nni", "-lterranni_hip",
          "-Wl,-rpath," + LIB_DIR, "-Wl,-rpath,$ORIGIN/../../../terranni_amd/_lib"], verbose)
    os.replace(exe + ".tmp", exe)
    tp = os.path.join(tdir, "tpie_test")
    _run([CXX, *CXX_FLAGS, *inc, os.path.join(ROOT, "tests", "cpp", "tpie_test.cpp"), "-o",
          tp + ".tmp"], verbose)
    os.replace(tp + ".tmp", tp)
    return out


def build_driver(verbose: bool = False) -> str:
    """terranni: the streaming DEM driver (LAS in, ESRI grids out) over
    libterranni_nni.so (plain engine) and libterranni_hip.so (batched)."""
    out = os.path.join(LIB_DIR, "terranni")
    _run([CXX, *CXX_FLAGS, "-fopenmp", "-I", os.path.join(ROOT, "include"),
          os.path.join(CSRC, "driver", "terranni.cpp"), "-o", out + ".tmp", "-L", LIB_DIR,
          "-lterranni_nni", "-lterranni_hip", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(out + ".tmp", out)
    return out


def build_stage_check(verbose: bool = False) -> str:
    """tests/cpp/_build/stage_bounds_test: a host program (no device code)
    that emulates the tile kernel's stage and label-tile indexing over the
    engine's own tile layout (prepare_tile, in libterranni_hip.so)."""
    tdir = os.path.join(ROOT, "tests", "cpp", "_build")
    os.makedirs(tdir, exist_ok=True)
    exe = os.path.join(tdir, "stage_bounds_test")
    _run([HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
          os.path.join(ROOT, "tests", "cpp", "stage_bounds_test.cpp"), "-o", exe + ".tmp",
          "-L", LIB_DIR, "-lterranni_hip", "-Wl,-rpath," + LIB_DIR,
          "-Wl,-rpath,$ORIGIN/../../../terranni_amd/_lib"], verbose)
    os.replace(exe + ".tmp", exe)
    return exe


def build_oracle(verbose: bool = False):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], verbose)


def build_all(verbose: bool = False):
    build_lib(verbose)
    build_nni_api(verbose)
    build_driver(verbose)
    build_stage_check(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-q" not in sys.argv)
