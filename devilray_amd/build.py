"""Build libdevilray.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m devilray_amd.build   (or devilray_amd.build.build_library()).
Objects go to devilray_amd/_build/, the shared library to devilray_amd/.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libdevilray.so")

SOURCES = ["api.hip", "sort.hip", "bvh_build.hip", "trace.hip", "light.hip", "render.hip", "guide.hip", "batch.hip",
           "display.hip", "instance.hip", "gen.cpp"]
HEADERS = ["dr_internal.hpp", "dr_device.hpp", "trace_device.hpp", "node_quant.hpp", "philox.hpp", "shade.hpp", "guide_device.hpp"]

# -ffp-contract=off: products and sums round as written, so the triangle test
# matches a CPU evaluation bit for bit; FMAs are written explicitly.
# The packed-FP32 feature is disabled: packed fp32 ops are avoided entirely.
HIPFLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return out


def build_library(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_host(verbose)
    return LIB


CPP = os.path.join(PKG, "cpp")
HOST_LIB = os.path.join(PKG, "libdevilray_cpp.so")
FACADE_TEST = os.path.join(CPP, "test_facade")
BENCH_CLI = os.path.join(CPP, "devil_ray_benchmark")  # the reference's benchmark command line


def build_host(verbose: bool = True) -> None:
    """The C++ facade (Renderer / Scene / cast over the C ABI) as
    libdevilray_cpp.so, and its test program; plain g++, no HIP headers."""
    # C++20: the facade keeps the reference's aggregates, designated
    # initialisers of materials and std::span views
    common = ["g++", "-std=c++20", "-O2", "-fPIC", "-Wall", "-Wextra", "-Wno-missing-field-initializers",
              f"-I{os.path.join(ROOT, 'include')}"]
    link = [f"-L{PKG}", "-ldevilray", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib"]
    steps = [
        common + ["-shared", "-o", HOST_LIB, os.path.join(CPP, "devilray.cpp")] + link,
        common + ["-o", FACADE_TEST, os.path.join(CPP, "test_facade.cpp"), f"-L{PKG}", "-ldevilray_cpp",
                  "-ldevilray", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib"],
        common + ["-O3", "-o", BENCH_CLI, os.path.join(CPP, "devil_ray_benchmark.cpp"), f"-L{PKG}",
                  "-ldevilray_cpp", "-ldevilray", "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib"],
    ]
    for cmd in steps:
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"g++ failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {HOST_LIB}, {FACADE_TEST} and {BENCH_CLI}")


if __name__ == "__main__":
    build_library()
    sys.exit(0)
