"""ctypes binding of libmpcd_gpu.so (the C ABI in include/mpcd_gpu.h).

The GPU library is the only compute path: loading raises if the library has
not been built, and every call raises ``MpcdError`` with the library's
message when it returns an error code (the C++ layer does the same with
std::runtime_error, as the reference does).
"""
from __future__ import annotations

import ctypes
import os
from typing import Dict, Optional

import numpy as np

from .params import MpcdParams

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libmpcd_gpu.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class MpcdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class WallInfo(ctypes.Structure):
    _fields_ = [("nBot", ctypes.c_int32), ("nTop", ctypes.c_int32), ("nLeft", ctypes.c_int32),
                ("nRight", ctypes.c_int32)] + [(k, ctypes.c_double) for k in (
                    "dEwall", "dPxBot", "dPxTop", "dPyBot", "dPyTop", "dPxLeft", "dPxRight", "dPyLeft", "dPyRight")]


class PassMetrics(ctypes.Structure):
    _fields_ = [("nZones", ctypes.c_int32), ("nParticlesMovedDense", ctypes.c_int32),
                ("nParticlesMovedSparse", ctypes.c_int32), ("_pad", ctypes.c_int32),
                ("corrRmsDU", ctypes.c_double), ("corrMaxDU", ctypes.c_double)]


CLOSURE_METRIC_NAMES = ("betaRepairOpt", "betaRepairApplied", "betaRepairNum", "betaRepairDen", "betaEOSApplied",
                        "meanPdriveRaw", "meanPdrive", "rmsRepairDU", "meanAbsRepairDU", "maxRepairDU", "rmsEOSDU",
                        "maxEOSDU", "momentumDeltaX", "momentumDeltaY")


class ClosureMetrics(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in CLOSURE_METRIC_NAMES] + [
        ("nMaskedInterzoneCells", ctypes.c_int32), ("_pad", ctypes.c_int32)]


FIELD_NAMES = ("Ux", "Uy", "Px", "Py", "T", "rho", "Pkin", "Pvir", "P", "rhoTarget")


class CellFieldsC(ctypes.Structure):
    _fields_ = [("N", _ip)] + [(k, _dp) for k in FIELD_NAMES]


class ClosureGrids(ctypes.Structure):
    _fields_ = [(k, _dp) for k in ("pdriveRaw", "pdrive", "duRepairX", "duRepairY", "duEOSX", "duEOSY")] + [
        ("interzoneMask", _ip), ("vRepaired", _dp)]


def _new_fields(Nc: int):
    """A CellFieldsC over fresh arrays, and the arrays (dict)."""
    arrs = {"N": np.zeros(Nc, np.int32)}
    for k in FIELD_NAMES:
        arrs[k] = np.zeros(Nc)
    return CellFieldsC(_ptr(arrs["N"], ctypes.c_int32), *[_ptr(arrs[k]) for k in FIELD_NAMES]), arrs


class StepMetrics(ctypes.Structure):
    _fields_ = [("wall", WallInfo), ("base", PassMetrics), ("shifted", PassMetrics), ("closure", ClosureMetrics)]


def struct_to_dict(s: ctypes.Structure) -> Dict[str, object]:
    out = {}
    for k, _ in s._fields_:
        if k.startswith("_"):
            continue
        val = getattr(s, k)
        out[k] = struct_to_dict(val) if isinstance(val, ctypes.Structure) else val
    return out


def _ptr(a: np.ndarray, ct=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _f64(a: np.ndarray) -> np.ndarray:
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"] or not a.flags["WRITEABLE"]:
        raise TypeError("expected a writable C-contiguous float64 array")
    return a


_LIB: Optional[ctypes.CDLL] = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the GPU library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it with __graft_entry__.build() "
                                "(or python -m src_incompressible_openmp_amd.build)")
    lib = ctypes.CDLL(path)
    lib.mpcd_gpu_last_error.restype = ctypes.c_char_p
    lib.mpcd_gpu_sim_create.argtypes = [ctypes.POINTER(MpcdParams), ctypes.POINTER(ctypes.c_void_p)]
    for name in ("mpcd_gpu_sim_destroy", "mpcd_gpu_sim_set_state", "mpcd_gpu_sim_get_state", "mpcd_gpu_sim_step",
                 "mpcd_gpu_sim_step_async", "mpcd_gpu_sim_check", "mpcd_gpu_sim_step_stages",
                 "mpcd_gpu_sim_synchronize", "mpcd_gpu_sim_stream"):
        getattr(lib, name).restype = ctypes.c_int
    _LIB = lib
    return lib


def _check(rc: int) -> None:
    if rc != 0:
        raise MpcdError(rc, load().mpcd_gpu_last_error().decode())


PROF_CLASSES = ("stream_bc", "bin", "collide", "cell_fields", "zone_pass", "zone_kernel", "closure", "kick",
                "zone_finish", "kick_stream")


def device_synchronize() -> None:
    _check(load().mpcd_gpu_device_synchronize())


def device_count() -> int:
    c = ctypes.c_int(0)
    _check(load().mpcd_gpu_device_count(ctypes.byref(c)))
    return c.value


# ---------------- stateless stage calls ----------------

def force_stream_bc(p: MpcdParams, x: np.ndarray, v: np.ndarray, seed: int, do_stream: bool = True):
    wi = WallInfo()
    _check(load().mpcd_gpu_force_stream_bc(ctypes.byref(p), _ptr(_f64(x)), _ptr(_f64(v)), ctypes.c_uint64(seed),
                                           int(do_stream), ctypes.byref(wi)))
    return struct_to_dict(wi)


def apply_bc_general(p: MpcdParams, x: np.ndarray, v: np.ndarray, seed: int):
    return force_stream_bc(p, x, v, seed, do_stream=False)


def srd_cell_id_with_random_shift(p: MpcdParams, x: np.ndarray, seed: int) -> np.ndarray:
    cid = np.zeros(p.n, np.int32)
    x = np.ascontiguousarray(x, np.float64)
    _check(load().mpcd_gpu_srd_cell_id_with_random_shift(ctypes.byref(p), _ptr(x), ctypes.c_uint64(seed),
                                                         _ptr(cid, ctypes.c_int32)))
    return cid


def srd_collision_step(p: MpcdParams, v: np.ndarray, cid: np.ndarray, seed: int) -> None:
    cid = np.ascontiguousarray(cid, np.int32)
    _check(load().mpcd_gpu_srd_collision_step(ctypes.byref(p), _ptr(_f64(v)), _ptr(cid, ctypes.c_int32),
                                              ctypes.c_uint64(seed)))


def compute_cell_fields(p: MpcdParams, x: np.ndarray, v: np.ndarray, rho_target: float) -> Dict[str, np.ndarray]:
    arrs = {"N": np.zeros(p.Nc, np.int32)}
    for k in FIELD_NAMES:
        arrs[k] = np.zeros(p.Nc)
    s = CellFieldsC(_ptr(arrs["N"], ctypes.c_int32), *[_ptr(arrs[k]) for k in FIELD_NAMES])
    x = np.ascontiguousarray(x, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    _check(load().mpcd_gpu_compute_cell_fields(ctypes.byref(p), _ptr(x), _ptr(v), ctypes.c_double(rho_target),
                                               ctypes.byref(s)))
    return arrs


def build_zone_layout(p: MpcdParams, mode: int):
    owner = np.zeros(p.Nc, np.int32)
    nz = ctypes.c_int32(0)
    _check(load().mpcd_gpu_build_zone_layout(ctypes.byref(p), mode, _ptr(owner, ctypes.c_int32), ctypes.byref(nz),
                                             None))
    rects = np.zeros(4 * nz.value, np.int32)
    _check(load().mpcd_gpu_build_zone_layout(ctypes.byref(p), mode, None, None, _ptr(rects, ctypes.c_int32)))
    return owner, rects.reshape(-1, 4)


def run_zone_kernel(p: MpcdParams, x: np.ndarray, v: np.ndarray, rect, seed: int, diag: bool = False):
    """run_zone_kernel_inplace on the rectangle (xs, xe, ys, ye), in place:
    returns (stats[6], moved[2]), with diag=True also ZoneKernelStats.diag[10]
    (reduced on the GPU)."""
    stats = np.zeros(6)
    moved = np.zeros(2, np.int32)
    dg = np.zeros(10)
    xs, xe, ys, ye = (int(t) for t in rect)
    _check(load().mpcd_gpu_run_zone_kernel(ctypes.byref(p), _ptr(_f64(x)), _ptr(_f64(v)), xs, xe, ys, ye,
                                           ctypes.c_uint64(seed), _ptr(stats), _ptr(moved, ctypes.c_int32),
                                           _ptr(dg) if diag else None))
    return (stats, moved, dg) if diag else (stats, moved)


def run_zone_kernel_mask(p: MpcdParams, x: np.ndarray, v: np.ndarray, mask, seed: int, diag: bool = False):
    """run_zone_kernel_inplace on the zone of any shape given by mask
    (ZoneDescriptor.maskVec: p.Nc entries, nonzero in the zone), in place;
    returns as run_zone_kernel."""
    m = np.ascontiguousarray(mask, np.uint8)
    if m.shape != (p.Nc,):
        raise ValueError("the zone mask needs one entry per cell")
    stats = np.zeros(6)
    moved = np.zeros(2, np.int32)
    dg = np.zeros(10)
    _check(load().mpcd_gpu_run_zone_kernel_mask(ctypes.byref(p), _ptr(_f64(x)), _ptr(_f64(v)),
                                                _ptr(m, ctypes.c_uint8), ctypes.c_uint64(seed), _ptr(stats),
                                                _ptr(moved, ctypes.c_int32), _ptr(dg) if diag else None))
    return (stats, moved, dg) if diag else (stats, moved)


def zone_pass(p: MpcdParams, x: np.ndarray, v: np.ndarray, layout: int, seed_base: int,
              gauss_seidel: bool = False):
    """Snapshot pass (run_zone_pass_openmp), or with gauss_seidel the zones
    applied one after another (run_zone_pass)."""
    m = PassMetrics()
    owner = np.zeros(p.Nc, np.int32)
    fn = load().mpcd_gpu_zone_pass_gauss_seidel if gauss_seidel else load().mpcd_gpu_zone_pass
    _check(fn(ctypes.byref(p), _ptr(_f64(x)), _ptr(_f64(v)), layout, ctypes.c_uint64(seed_base), ctypes.byref(m),
              _ptr(owner, ctypes.c_int32)))
    return struct_to_dict(m), owner


class OccupancyStats(ctypes.Structure):
    _fields_ = [("occStd", ctypes.c_double), ("outBand", ctypes.c_double), ("nDense", ctypes.c_int32),
                ("nSparse", ctypes.c_int32), ("nEmptyActive", ctypes.c_int32), ("_pad", ctypes.c_int32)]


class StepDiag(ctypes.Structure):
    _fields_ = [("base", OccupancyStats), ("shifted", OccupancyStats)] + [
        (k, ctypes.c_double) for k in ("meanKinetic", "meanVelocityX", "meanVelocityY", "meanUx", "meanUy", "Qx")]


def occupancy_stats(p: MpcdParams, counts: np.ndarray) -> Dict[str, float]:
    """occStd, outBand, nDense, nSparse, nEmptyActive of a zone pass's before
    or after counts (reduced on the GPU)."""
    c = np.ascontiguousarray(counts, np.int32)
    if c.size != p.Nc:
        raise ValueError("counts must have Nc entries")
    out = OccupancyStats()
    _check(load().mpcd_gpu_occupancy_stats(ctypes.byref(p), _ptr(c, ctypes.c_int32), ctypes.byref(out)))
    return struct_to_dict(out)


def closure_training_pair(p: MpcdParams, ref: Dict[str, np.ndarray], grids: Dict[str, np.ndarray], beta: float,
                          out_scale: float):
    """The surrogate's (features [Nc,4], target [Nc,2]) of one closure
    snapshot (reference CellFields N, Ux, Uy, T; grids duRepairX/Y,
    duEOSX/Y; the applied beta), computed on the GPU."""
    Nc = p.Nc
    f = np.zeros((Nc, 4), np.float32)
    t = np.zeros((Nc, 2), np.float32)
    a = [np.ascontiguousarray(ref["N"], np.int32)] + [np.ascontiguousarray(ref[k], np.float64)
                                                       for k in ("Ux", "Uy", "T")]
    g = [np.ascontiguousarray(grids[k], np.float64) for k in ("duRepairX", "duRepairY", "duEOSX", "duEOSY")]
    if any(q.size != Nc for q in a + g):
        raise ValueError("every field and grid must have Nc entries")
    _check(load().mpcd_gpu_closure_training_pair(
        ctypes.byref(p), _ptr(a[0], ctypes.c_int32), *[_ptr(q) for q in a[1:]], *[_ptr(q) for q in g],
        ctypes.c_double(beta), ctypes.c_double(out_scale), f.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
        t.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
    return f, t


def liquid_closure(p: MpcdParams, x_ref, v_ref, x_red, v_red):
    n2 = 2 * p.n
    xo, vo, vrep = np.zeros(n2), np.zeros(n2), np.zeros(n2)
    m = ClosureMetrics()
    grids = {k: np.zeros(p.Nc) for k in ("pdriveRaw", "pdrive", "duRepairX", "duRepairY", "duEOSX", "duEOSY")}
    mask = np.zeros(p.Nc, np.int32)
    g = ClosureGrids(*[_ptr(grids[k]) for k in ("pdriveRaw", "pdrive", "duRepairX", "duRepairY", "duEOSX",
                                                "duEOSY")], _ptr(mask, ctypes.c_int32), _ptr(vrep))
    args = [np.ascontiguousarray(a, np.float64) for a in (x_ref, v_ref, x_red, v_red)]
    _check(load().mpcd_gpu_liquid_closure(ctypes.byref(p), *[_ptr(a) for a in args], _ptr(xo), _ptr(vo),
                                          ctypes.byref(m), ctypes.byref(g)))
    out = {"x": xo, "v": vo, "v_repaired": vrep, "metrics": struct_to_dict(m), "mask": mask}
    out.update(grids)
    return out


# ---------------- resident simulation ----------------

class Sim:
    """Device-resident SRD/MPCD state advanced by whole time steps."""

    def __init__(self, p: MpcdParams):
        self.p = p
        self._h = ctypes.c_void_p()
        _check(load().mpcd_gpu_sim_create(ctypes.byref(p), ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h:
            _check(load().mpcd_gpu_sim_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, x, v) -> None:
        """Uploads positions and velocities (interleaved, particle-id order);
        either may be None to keep each particle's current value."""
        x = None if x is None else np.ascontiguousarray(x, np.float64)
        v = None if v is None else np.ascontiguousarray(v, np.float64)
        for a in (x, v):
            if a is not None and a.size != 2 * self.p.n:
                raise ValueError("state size does not match params.n")
        _check(load().mpcd_gpu_sim_set_state(self._h, None if x is None else _ptr(x), None if v is None else _ptr(v)))

    def get_state(self):
        x, v = np.zeros(2 * self.p.n), np.zeros(2 * self.p.n)
        _check(load().mpcd_gpu_sim_get_state(self._h, _ptr(x), _ptr(v)))
        return x, v

    def step(self, first_step: int, n_steps: int = 1, metrics: bool = False):
        m = StepMetrics() if metrics else None
        _check(load().mpcd_gpu_sim_step(self._h, first_step, n_steps, ctypes.byref(m) if m is not None else None))
        return struct_to_dict(m) if m is not None else None

    def step_async(self, first_step: int, n_steps: int = 1) -> None:
        _check(load().mpcd_gpu_sim_step_async(self._h, first_step, n_steps))

    def check(self) -> None:
        _check(load().mpcd_gpu_sim_check(self._h))

    def synchronize(self) -> None:
        _check(load().mpcd_gpu_sim_synchronize(self._h))

    def profile(self, enable: bool) -> None:
        """Start (True) or stop per-class kernel timing; starting resets it."""
        _check(load().mpcd_gpu_sim_profile(self._h, int(enable)))

    def profile_read(self):
        ms = np.zeros(len(PROF_CLASSES))
        cnt = np.zeros(len(PROF_CLASSES), np.int32)
        _check(load().mpcd_gpu_sim_profile_read(self._h, _ptr(ms), _ptr(cnt, ctypes.c_int32)))
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(PROF_CLASSES)}

    def set_surrogate(self, net, K: int, out_scale: float, du_cap: float) -> None:
        """Steps s with s % K != 0 use the closure surrogate `net`
        (src_incompressible_openmp_amd.surrogate.Net); net None or K <= 0:
        plain closure every step."""
        h = net._h if net is not None else ctypes.c_void_p()
        _check(load().mpcd_gpu_sim_set_surrogate(self._h, h, int(K), ctypes.c_double(out_scale),
                                                 ctypes.c_double(du_cap)))

    def set_surrogate_check(self, occ_tol: float, div_tol: float) -> None:
        """Tolerances of the surrogate's residual check (occupancy spread
        and divergence residual relative to the last plain step)."""
        _check(load().mpcd_gpu_sim_set_surrogate_check(self._h, ctypes.c_double(occ_tol), ctypes.c_double(div_tol)))

    def row_profile(self, enable: int = -1, fetch: bool = False):
        """Row profiles summed on the device after each step: enable 1 clears
        and starts, 0 stops, -1 keeps; with fetch returns (sums of vx, sums of
        |v|^2, particle counts) per cell row and the number of steps."""
        rows = np.zeros(3 * self.p.Ny) if fetch else None
        st = ctypes.c_int32(0)
        _check(load().mpcd_gpu_sim_row_profile(self._h, int(enable), _ptr(rows) if fetch else None,
                                               ctypes.byref(st) if fetch else None))
        if not fetch:
            return None
        Ny = self.p.Ny
        return rows[:Ny], rows[Ny:2 * Ny], rows[2 * Ny:], int(st.value)

    def surrogate_counts(self):
        c = np.zeros(3, np.int32)
        _check(load().mpcd_gpu_sim_surrogate_counts(self._h, _ptr(c, ctypes.c_int32)))
        return {"learned": int(c[0]), "plain": int(c[1]), "fallback": int(c[2])}

    def closure_sample(self, enable: bool = True, out_scale: float = 0.0, fetch: bool = False):
        """Enables storing each plain closure's features and target
        (correction / out_scale); with fetch, returns the last pair."""
        f = np.zeros((self.p.Nc, 4), np.float32) if fetch else None
        t = np.zeros((self.p.Nc, 2), np.float32) if fetch else None
        _check(load().mpcd_gpu_sim_closure_sample(
            self._h, int(enable), ctypes.c_double(out_scale),
            f.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if fetch else None,
            t.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if fetch else None))
        return (f, t) if fetch else None

    def closure_snapshot(self, enable: bool = True, fetch: bool = False):
        """Enables keeping each plain closure's input and output (the
        reference and red CellFields, the six closure grids, beta); with
        fetch, returns the last one kept as a dict (None if none yet)."""
        if not fetch:
            _check(load().mpcd_gpu_sim_closure_snapshot(self._h, int(enable), None, None, None, None, None))
            return None
        Nc = self.p.Nc
        (sref, ref), (sred, red) = _new_fields(Nc), _new_fields(Nc)
        names = ("pdriveRaw", "pdrive", "duRepairX", "duRepairY", "duEOSX", "duEOSY")
        grids = {k: np.zeros(Nc) for k in names}
        grids["interzoneMask"] = np.zeros(Nc, np.int32)
        cg = ClosureGrids(*[_ptr(grids[k]) for k in names], _ptr(grids["interzoneMask"], ctypes.c_int32))
        beta, step = ctypes.c_double(0.0), ctypes.c_int32(-1)
        _check(load().mpcd_gpu_sim_closure_snapshot(self._h, int(enable), ctypes.byref(sref), ctypes.byref(sred),
                                                     ctypes.byref(cg), ctypes.byref(beta), ctypes.byref(step)))
        if step.value < 0:
            return None
        return {"step": step.value, "betaRepairApplied": beta.value, "ref": ref, "red": red, **grids}

    def stream(self) -> int:
        s = ctypes.c_void_p()
        _check(load().mpcd_gpu_sim_stream(self._h, ctypes.byref(s)))
        return s.value or 0

    def step_diag(self, step: int):
        """One step with its metrics and its device-reduced diagnostics
        (occupancy after each pass, means of the output state and of its
        cell fields): returns (metrics, diag) dictionaries."""
        m, d = StepMetrics(), StepDiag()
        _check(load().mpcd_gpu_sim_step_diag(self._h, step, ctypes.byref(m), ctypes.byref(d)))
        return struct_to_dict(m), struct_to_dict(d)

    def cell_fields(self, rho_target: float):
        """compute_cell_fields of the resident state: dict of Nc arrays."""
        Nc = self.p.Nc
        arrs = {"N": np.zeros(Nc, np.int32)}
        for k in FIELD_NAMES:
            arrs[k] = np.zeros(Nc)
        f = CellFieldsC(_ptr(arrs["N"], ctypes.c_int32), *[_ptr(arrs[k]) for k in FIELD_NAMES])
        _check(load().mpcd_gpu_sim_cell_fields(self._h, ctypes.c_double(rho_target), ctypes.byref(f)))
        return arrs

    def step_stages(self, step: int):
        n2 = 2 * self.p.n
        st = {k: np.zeros(n2) for k in ("x_ref", "v_ref", "x_base", "v_base", "x_red", "v_red")}
        m = StepMetrics()
        _check(load().mpcd_gpu_sim_step_stages(self._h, step, *[_ptr(st[k]) for k in (
            "x_ref", "v_ref", "x_base", "v_base", "x_red", "v_red")], ctypes.byref(m)))
        st["metrics"] = struct_to_dict(m)
        return st

    def set_seed_key(self, key: int) -> None:
        """Every following step draws with the benchmark driver's seeds plus
        key * 0x9E3779B97F4A7C15 (key 0: the driver's own)."""
        _check(load().mpcd_gpu_sim_set_seed_key(self._h, ctypes.c_uint64(int(key))))


# ---------------- batched ensembles ----------------

class Ensemble:
    """Independent channels (replicas with their own parameters and seed
    keys) advanced together on the device: every stage of a step is one
    launch over all replicas, and the step can be replayed as a HIP graph.
    Replica r follows a Sim of params[r] with set_seed_key(keys[r]) bit for
    bit."""

    def __init__(self, params, seed_keys=None):
        self.params = list(params)
        R = len(self.params)
        arr = (MpcdParams * R)(*self.params)
        keys = None
        if seed_keys is not None:
            if len(seed_keys) != R:
                raise ValueError("one seed key per replica")
            keys = (ctypes.c_uint64 * R)(*[int(k) for k in seed_keys])
        self._h = ctypes.c_void_p()
        _check(load().mpcd_gpu_ensemble_create(arr, keys, R, ctypes.byref(self._h)))

    def close(self) -> None:
        if self._h:
            _check(load().mpcd_gpu_ensemble_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        n, c = ctypes.c_int64(), ctypes.c_int32()
        z = (ctypes.c_int32 * 2)()
        _check(load().mpcd_gpu_ensemble_info(self._h, ctypes.byref(n), ctypes.byref(c), z))
        return {"particles": n.value, "cells": c.value, "zones": [z[0], z[1]]}

    def set_state(self, r: int, x, v) -> None:
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        if x.size != 2 * self.params[r].n or v.size != 2 * self.params[r].n:
            raise ValueError("state size does not match the replica's n")
        _check(load().mpcd_gpu_ensemble_set_state(self._h, int(r), _ptr(x), _ptr(v)))

    def get_state(self, r: int):
        n2 = 2 * self.params[r].n
        x, v = np.zeros(n2), np.zeros(n2)
        _check(load().mpcd_gpu_ensemble_get_state(self._h, int(r), _ptr(x), _ptr(v)))
        return x, v

    def step(self, first_step: int, n_steps: int = 1, graph: bool = True, wait: bool = True) -> None:
        _check(load().mpcd_gpu_ensemble_step(self._h, int(first_step), int(n_steps), int(graph)))
        if wait:
            self.synchronize()

    def synchronize(self) -> None:
        _check(load().mpcd_gpu_ensemble_synchronize(self._h))

    def flow(self, enable: bool = True, fetch: bool = False):
        """Enables (clearing) or disables the per-replica Qx sums; with
        fetch, returns (sum Qx, sum Qx^2, steps) so far."""
        R = len(self.params)
        if not fetch:
            _check(load().mpcd_gpu_ensemble_flow(self._h, int(enable), None, None, None))
            return None
        q, q2 = np.zeros(R), np.zeros(R)
        ns = ctypes.c_int32()
        _check(load().mpcd_gpu_ensemble_flow(self._h, int(enable), _ptr(q), _ptr(q2), ctypes.byref(ns)))
        return q, q2, ns.value
