"""Build libmpcd_gpu.so (gfx950) in-tree with hipcc.

FP contraction is off so every FP64 product and sum rounds as written (the
parity contract with the reference is bitwise for positions and cell
decisions); packed FP32 instructions are disabled on every HIP compile.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libmpcd_gpu.so")

HIP_SOURCES = [os.path.join(CSRC, "mpcd_engine.hip"), os.path.join(CSRC, "mpcd_net.hip")]
CXX_SOURCES = [os.path.join(CSRC, "mpcd_api.cpp"), os.path.join(CSRC, "host_io.cpp")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"),
          "-I" + os.path.join(ROOT, "include", "mpcd"), "-I" + CSRC]
# -simplifycfg-sink-common=false: merging the two arms of a branch that write
# through different fields of a kernel's per-lane context turns the field
# choice into a computed offset, which keeps the context in private memory and
# makes the zone kernel's working-set accesses generic (flat) ones.
HIP_FLAGS = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-mllvm", "-simplifycfg-sink-common=false"]


DRIVER_SRC = os.path.join(CSRC, "main_benchmark_poiseuille_gpu.cpp")
DRIVER = os.path.join(PKG, "main_benchmark_poiseuille_gpu")


def build_driver(verbose: bool = False) -> str:
    """The benchmark driver on the resident simulation, linked against the
    library next to it."""
    cmd = ["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "include", "mpcd"),
           DRIVER_SRC, "-L" + PKG, "-lmpcd_gpu", "-Wl,-rpath,$ORIGIN", "-o", DRIVER + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("g++ failed building main_benchmark_poiseuille_gpu")
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


def build(verbose: bool = False) -> str:
    sources = HIP_SOURCES + [s for s in CXX_SOURCES if os.path.exists(s)]
    cmd = [HIPCC, *HIP_FLAGS, *COMMON, "-shared", *sources, "-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libmpcd_gpu.so")
    os.replace(LIB + ".tmp", LIB)
    build_driver(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
