"""Reference file formats (byte-identical) and the initial-state generator.

* state binaries: headerless little-endian FP64 interleaved ``[x1,y1,...]``
  (x, v, r0) and u8 type (state_io.cpp:9-68);
* ``cellfields.bin``: int32 Nx, Ny, int32 N[Nc], then 10 FP64 arrays
  Ux, Uy, Px, Py, T, rho, Pkin, Pvir, P, rhoTarget (dump_io.cpp:19-39);
* i32 / f64 grids: int32 Nx, Ny then Nc values, cell index ``iy + Ny*ix``
  (zone_pass.cpp:1165-1181, liquid_closure.cpp:446-480).
"""
from __future__ import annotations

from typing import Dict, Tuple

import numpy as np

CELL_FIELD_NAMES = ("Ux", "Uy", "Px", "Py", "T", "rho", "Pkin", "Pvir", "P", "rhoTarget")


def read_xy(path: str, n: int) -> np.ndarray:
    a = np.fromfile(path, dtype="<f8", count=2 * n)
    if a.size != 2 * n:
        raise RuntimeError(f"Cannot read expected payload from: {path}")
    return a


def write_xy(path: str, a: np.ndarray, n: int) -> None:
    a = np.ascontiguousarray(a, dtype="<f8")
    if a.size != 2 * n:
        raise RuntimeError(f"Interleaved XY buffer has unexpected size for write: {path}")
    a.tofile(path)


def read_u8(path: str, n: int) -> np.ndarray:
    a = np.fromfile(path, dtype=np.uint8, count=n)
    if a.size != n:
        raise RuntimeError(f"Cannot read expected payload from: {path}")
    return a


def write_u8(path: str, a: np.ndarray, n: int) -> None:
    a = np.ascontiguousarray(a, dtype=np.uint8)
    if a.size != n:
        raise RuntimeError(f"u8 buffer has unexpected size for write: {path}")
    a.tofile(path)


def write_cell_fields(path: str, Nx: int, Ny: int, fields: Dict[str, np.ndarray]) -> None:
    with open(path, "wb") as fh:
        np.array([Nx, Ny], dtype="<i4").tofile(fh)
        np.ascontiguousarray(fields["N"], dtype="<i4").tofile(fh)
        for k in CELL_FIELD_NAMES:
            np.ascontiguousarray(fields[k], dtype="<f8").tofile(fh)


def read_cell_fields(path: str) -> Dict[str, np.ndarray]:
    raw = np.fromfile(path, dtype=np.uint8)
    Nx, Ny = np.frombuffer(raw[:8].tobytes(), dtype="<i4")
    Nc = int(Nx) * int(Ny)
    out: Dict[str, np.ndarray] = {"Nx": int(Nx), "Ny": int(Ny)}
    off = 8
    out["N"] = np.frombuffer(raw[off:off + 4 * Nc].tobytes(), dtype="<i4").copy()
    off += 4 * Nc
    for k in CELL_FIELD_NAMES:
        out[k] = np.frombuffer(raw[off:off + 8 * Nc].tobytes(), dtype="<f8").copy()
        off += 8 * Nc
    if off != raw.size:
        raise RuntimeError(f"cellfields payload size mismatch in {path}")
    return out


def write_grid(path: str, Nx: int, Ny: int, data: np.ndarray, dtype: str) -> None:
    data = np.ascontiguousarray(data, dtype=dtype)
    if data.size != Nx * Ny:
        raise RuntimeError("grid: unexpected data size")
    with open(path, "wb") as fh:
        np.array([Nx, Ny], dtype="<i4").tofile(fh)
        data.tofile(fh)


def read_grid(path: str, dtype: str) -> Tuple[int, int, np.ndarray]:
    raw = np.fromfile(path, dtype=np.uint8)
    Nx, Ny = np.frombuffer(raw[:8].tobytes(), dtype="<i4")
    data = np.frombuffer(raw[8:].tobytes(), dtype=dtype).copy()
    if data.size != int(Nx) * int(Ny):
        raise RuntimeError(f"grid payload size mismatch in {path}")
    return int(Nx), int(Ny), data


def uniform_state(n: int, Lx: float, Ly: float, kBT: float = 1.0, seed: int = 1):
    """Initial channel state: positions uniform in [0,Lx)x[0,Ly), velocities
    N(0, kBT) per component minus the mean, r0 = x, type = 0
    (build_ad_hoc_state_uniform.m:27-39).  numpy PCG64 stream from ``seed``."""
    rng = np.random.default_rng(seed)
    x = np.empty(2 * n)
    x[0::2] = Lx * rng.random(n)
    x[1::2] = Ly * rng.random(n)
    v = np.sqrt(kBT) * rng.standard_normal(2 * n)
    v[0::2] -= v[0::2].mean()
    v[1::2] -= v[1::2].mean()
    return x, v, np.zeros(n, dtype=np.uint8), x.copy()
