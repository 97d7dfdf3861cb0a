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
        os.path.join(ROOT, "src_incompressible_openmp_amd", "csrc", "mpcd_zone.h"),
        os.path.join(ROOT, "src_incompressible_openmp_amd", "csrc", "mpcd_host.h"),
        os.path.join(ROOT, "include", "mpcd_rng.h"), os.path.join(ROOT, "include", "mpcd_gpu.h")]

_dp = ctypes.POINTER(ctypes.c_double)


SO_SMALL = os.path.join(OUT_DIR, "libcore_host_small.so")
# staging limits small enough that the golden cases run past them
SMALL_DEFS = ["-DMPCD_ZONE_MAX_MOVE=2", "-DMPCD_ZONE_TMP_L=4"]


def build(small: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    so = SO_SMALL if small else SO
    if not os.path.exists(so) or any(os.path.getmtime(d) > os.path.getmtime(so) for d in DEPS):
        subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off",
                        *(SMALL_DEFS if small else []),
                        "-I" + os.path.join(ROOT, "include"),
                        "-I" + os.path.join(ROOT, "src_incompressible_openmp_amd", "csrc"), SRC, "-o", so],
                       check=True)
    return so


class CoreHost:
    def __init__(self, small: bool = False):
        self.lib = ctypes.CDLL(build(small))

    @staticmethod
    def _p(a, ct=ctypes.c_double):
        return a.ctypes.data_as(ctypes.POINTER(ct))

    def force_stream_bc(self, p, x, v, seed, do_stream=True):
        return self.lib.core_force_stream_bc(ctypes.byref(p), self._p(x), self._p(v), ctypes.c_uint64(seed),
                                             int(do_stream))

    def collide(self, p, v, cid, seed):
        cid = np.ascontiguousarray(cid, np.int32)
        return self.lib.core_collide(ctypes.byref(p), self._p(v), self._p(cid, ctypes.c_int32), ctypes.c_uint64(seed))

    def zone_pass(self, p, x, v, layout, seed_base, wide=False, ws_bytes=38 * 1024):
        """wide: int working set (else uint16_t); ws_bytes: working-set bytes,
        list links beyond them go to the zone's global region."""
        moved = np.zeros(2, np.int32)
        corr = np.zeros(2)
        rc = self.lib.core_zone_pass(ctypes.byref(p), self._p(x), self._p(v), layout, ctypes.c_uint64(seed_base),
                                     self._p(moved, ctypes.c_int32), self._p(corr), int(wide),
                                     ctypes.c_longlong(ws_bytes))
        if rc != 0:
            raise RuntimeError(f"core_zone_pass failed: {rc}")
        return moved, corr
