"""ctypes binding of the CPU oracle (oracle/liboracle.so).

The oracle is the checker: tests call it, the product never does.  mode 0 =
libstdc++ draws (the reference's own stream), mode 1 = counter-based draws
(the contract of the GPU library).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Dict, Optional

import numpy as np

from src_incompressible_openmp_amd.params import MpcdParams

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")

STD, CTR = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


from src_incompressible_openmp_amd.gpu import (CLOSURE_METRIC_NAMES, FIELD_NAMES, CellFieldsC,  # noqa: E402
                                              ClosureMetrics, PassMetrics, WallInfo, struct_to_dict)


def _ptr(a: np.ndarray, ct=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def new_fields(Nc: int):
    arrs = {"N": np.zeros(Nc, np.int32)}
    for k in FIELD_NAMES:
        arrs[k] = np.zeros(Nc)
    s = CellFieldsC(_ptr(arrs["N"], ctypes.c_int32), *[_ptr(arrs[k]) for k in FIELD_NAMES])
    return s, arrs


def build(quiet: bool = True) -> None:
    subprocess.run(["make", "-C", ORACLE_DIR, "-j8"], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


class Oracle:
    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build()
        self.lib = ctypes.CDLL(path)
        L = self.lib
        L.oracle_last_error.restype = ctypes.c_char_p
        L.oracle_ctr_log.restype = ctypes.c_double
        L.oracle_ctr_log.argtypes = [ctypes.c_double]
        for name in ("oracle_apply_bc", "oracle_shift_cell_ids", "oracle_collide", "oracle_zone_kernel",
                     "oracle_zone_pass", "oracle_step"):
            getattr(L, name).restype = ctypes.c_int

    def _check(self, rc: int) -> None:
        if rc != 0:
            raise RuntimeError(self.lib.oracle_last_error().decode())

    # ---- stages ----
    def force_stream(self, p: MpcdParams, x: np.ndarray, v: np.ndarray) -> None:
        self._check(self.lib.oracle_force_stream(ctypes.byref(p), _ptr(x), _ptr(v)))

    def apply_bc(self, mode: int, p: MpcdParams, x: np.ndarray, v: np.ndarray, seed: int) -> Dict[str, float]:
        wi = WallInfo()
        self._check(self.lib.oracle_apply_bc(mode, ctypes.byref(p), _ptr(x), _ptr(v), ctypes.c_uint64(seed),
                                             ctypes.byref(wi)))
        return struct_to_dict(wi)

    def shift_cell_ids(self, mode: int, p: MpcdParams, x: np.ndarray, seed: int):
        cid = np.zeros(p.n, np.int32)
        sh = np.zeros(2)
        self._check(self.lib.oracle_shift_cell_ids(mode, ctypes.byref(p), _ptr(x), ctypes.c_uint64(seed),
                                                   _ptr(cid, ctypes.c_int32), _ptr(sh)))
        return cid, (sh[0], sh[1])

    def collide(self, mode: int, p: MpcdParams, v: np.ndarray, cid: np.ndarray, seed: int) -> None:
        cid = np.ascontiguousarray(cid, np.int32)
        self._check(self.lib.oracle_collide(mode, ctypes.byref(p), _ptr(v), _ptr(cid, ctypes.c_int32),
                                            ctypes.c_uint64(seed)))

    def cell_fields(self, p: MpcdParams, x: np.ndarray, v: np.ndarray, rho_target: float):
        s, arrs = new_fields(p.Nc)
        self._check(self.lib.oracle_cell_fields(ctypes.byref(p), _ptr(x), _ptr(v), ctypes.c_double(rho_target),
                                                ctypes.byref(s)))
        return arrs

    def layout(self, p: MpcdParams, mode: int):
        owner = np.zeros(p.Nc, np.int32)
        nz = ctypes.c_int32()
        self._check(self.lib.oracle_layout(ctypes.byref(p), mode, _ptr(owner, ctypes.c_int32), ctypes.byref(nz)))
        return owner, nz.value

    def zone_kernel(self, mode: int, p: MpcdParams, x, v, layout: int, zone_index: int, seed: int):
        st = np.zeros(6)
        dg = np.zeros(10)
        mv = np.zeros(2, np.int32)
        self._check(self.lib.oracle_zone_kernel(mode, ctypes.byref(p), _ptr(x), _ptr(v), layout, zone_index,
                                                ctypes.c_uint64(seed), _ptr(st), _ptr(dg), _ptr(mv, ctypes.c_int32)))
        return st, dg, mv

    def zone_kernel_mask(self, mode: int, p: MpcdParams, x, v, mask, seed: int):
        """run_zone_kernel_inplace on the zone given by mask (p.Nc bytes)"""
        st = np.zeros(6)
        dg = np.zeros(10)
        mv = np.zeros(2, np.int32)
        m = np.ascontiguousarray(mask, np.uint8)
        self._check(self.lib.oracle_zone_kernel_mask(mode, ctypes.byref(p), _ptr(x), _ptr(v), _ptr(m, ctypes.c_uint8),
                                                     ctypes.c_uint64(seed), _ptr(st), _ptr(dg),
                                                     _ptr(mv, ctypes.c_int32)))
        return st, dg, mv

    def zone_pass(self, mode: int, p: MpcdParams, x, v, layout: int, seed_base: int, gauss_seidel: bool = False):
        m = PassMetrics()
        owner = np.zeros(p.Nc, np.int32)
        self._check(self.lib.oracle_zone_pass(mode, ctypes.byref(p), _ptr(x), _ptr(v), layout,
                                              ctypes.c_uint64(seed_base), int(gauss_seidel), ctypes.byref(m),
                                              _ptr(owner, ctypes.c_int32)))
        return struct_to_dict(m), owner

    def zone_pass_full(self, mode: int, p: MpcdParams, x, v, layout: int, seed_base: int):
        """The snapshot pass with a private copy of the domain per zone (the
        reference's structure), in place; returns the pass metrics."""
        m = PassMetrics()
        self._check(self.lib.oracle_zone_pass_full(mode, ctypes.byref(p), _ptr(x), _ptr(v), layout,
                                                   ctypes.c_uint64(seed_base), ctypes.byref(m)))
        return struct_to_dict(m)

    def closure(self, p: MpcdParams, x_ref, v_ref, x_red, v_red):
        n2 = 2 * p.n
        xo, vo, vrep = np.zeros(n2), np.zeros(n2), np.zeros(n2)
        m = ClosureMetrics()
        grids = [np.zeros(p.Nc) for _ in range(6)]
        gp = (_dp * 6)(*[_ptr(g) for g in grids])
        mask = np.zeros(p.Nc, np.int32)
        fs = (CellFieldsC * 4)()
        farrs = []
        for k in range(4):
            s, arrs = new_fields(p.Nc)
            fs[k] = s
            farrs.append(arrs)
        self._check(self.lib.oracle_closure(ctypes.byref(p), _ptr(x_ref), _ptr(v_ref), _ptr(x_red), _ptr(v_red),
                                            _ptr(xo), _ptr(vo), ctypes.byref(m), gp, _ptr(mask, ctypes.c_int32), fs,
                                            _ptr(vrep)))
        names = ("pdriveRaw", "pdrive", "duRepairX", "duRepairY", "duEOSX", "duEOSY")
        out = {"x": xo, "v": vo, "v_repaired": vrep, "metrics": struct_to_dict(m), "mask": mask,
               "fields": dict(zip(("ref", "red", "repaired", "out"), farrs))}
        out.update(dict(zip(names, grids)))
        return out

    def step(self, mode: int, p: MpcdParams, x, v, step: int, stages: bool = False):
        n2 = 2 * p.n
        st = {k: np.zeros(n2) for k in ("x_ref", "v_ref", "x_base", "v_base", "x_red", "v_red")} if stages else {}
        mb, ms, mc, wi = PassMetrics(), PassMetrics(), ClosureMetrics(), WallInfo()
        args = [(_ptr(st[k]) if stages else None) for k in ("x_ref", "v_ref", "x_base", "v_base", "x_red", "v_red")]
        self._check(self.lib.oracle_step(mode, ctypes.byref(p), _ptr(x), _ptr(v), step, *args, ctypes.byref(mb),
                                         ctypes.byref(ms), ctypes.byref(mc), ctypes.byref(wi)))
        st.update({"base": struct_to_dict(mb), "shifted": struct_to_dict(ms), "closure": struct_to_dict(mc),
                   "wall": struct_to_dict(wi)})
        return st

    def occupancy_stats(self, p: MpcdParams, counts: np.ndarray):
        from src_incompressible_openmp_amd.gpu import OccupancyStats

        c = np.ascontiguousarray(counts, np.int32)
        out = OccupancyStats()
        self._check(self.lib.oracle_occupancy_stats(ctypes.byref(p), _ptr(c, ctypes.c_int32), ctypes.byref(out)))
        return struct_to_dict(out)

    def interzone_mask(self, p: MpcdParams) -> np.ndarray:
        out = np.zeros(p.Nc, np.int32)
        self._check(self.lib.oracle_interzone_mask(ctypes.byref(p), _ptr(out, ctypes.c_int32)))
        return out

    # ---- counter generator internals ----
    def ctr_philox(self, seed: int, ctr) -> np.ndarray:
        c = np.asarray(ctr, np.uint32)
        out = np.zeros(4, np.uint32)
        self.lib.oracle_ctr_philox(ctypes.c_uint64(seed), _ptr(c, ctypes.c_uint32), _ptr(out, ctypes.c_uint32))
        return out

    def ctr_log(self, x: float) -> float:
        return self.lib.oracle_ctr_log(ctypes.c_double(x))

    def ctr_sincos_2pi(self, u: float):
        sc = np.zeros(2)
        self.lib.oracle_ctr_sincos_2pi(ctypes.c_double(u), _ptr(sc))
        return sc[0], sc[1]

    def ctr_normal_pair(self, seed: int, pid: int, sweep: int, wall: int):
        out = np.zeros(2)
        self.lib.oracle_ctr_normal_pair(ctypes.c_uint64(seed), ctypes.c_uint32(pid), sweep, wall, _ptr(out))
        return out[0], out[1]


_ORACLE: Optional[Oracle] = None


def get_oracle() -> Oracle:
    global _ORACLE
    if _ORACLE is None:
        _ORACLE = Oracle()
    return _ORACLE
