"""Build and bind tests/core_host.cpp: the GPU library's kernel bodies
compiled for the host (test infrastructure for the host logic check)."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "core_host.cpp")
OUT_DIR = os.path.join(ROOT, "tests", "_build")
SO = os.path.join(OUT_DIR, "libcore_host.so")
DEPS = [SRC, os.path.join(ROOT, "src_incompressible_openmp_amd", "csrc", "mpcd_core.h"),
        os.path.join(ROOT, "src_incompressible_openmp_amd", "csrc", "mpcd_host.h"),
        os.path.join(ROOT, "include", "mpcd_rng.h"), os.path.join(ROOT, "include", "mpcd_gpu.h")]

_dp = ctypes.POINTER(ctypes.c_double)


def build() -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    if not os.path.exists(SO) or any(os.path.getmtime(d) > os.path.getmtime(SO) for d in DEPS):
        subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off",
                        "-I" + os.path.join(ROOT, "include"),
                        "-I" + os.path.join(ROOT, "src_incompressible_openmp_amd", "csrc"), SRC, "-o", SO],
                       check=True)
    return SO


class CoreHost:
    def __init__(self):
        self.lib = ctypes.CDLL(build())

    @staticmethod
    def _p(a, ct=ctypes.c_double):
        return a.ctypes.data_as(ctypes.POINTER(ct))

    def force_stream_bc(self, p, x, v, seed, do_stream=True):
        return self.lib.core_force_stream_bc(ctypes.byref(p), self._p(x), self._p(v), ctypes.c_uint64(seed),
                                             int(do_stream))

    def collide(self, p, v, cid, seed):
        cid = np.ascontiguousarray(cid, np.int32)
        return self.lib.core_collide(ctypes.byref(p), self._p(v), self._p(cid, ctypes.c_int32), ctypes.c_uint64(seed))

    def zone_pass(self, p, x, v, layout, seed_base):
        moved = np.zeros(2, np.int32)
        corr = np.zeros(2)
        rc = self.lib.core_zone_pass(ctypes.byref(p), self._p(x), self._p(v), layout, ctypes.c_uint64(seed_base),
                                     self._p(moved, ctypes.c_int32), self._p(corr))
        if rc != 0:
            raise RuntimeError(f"core_zone_pass failed: {rc}")
        return moved, corr
