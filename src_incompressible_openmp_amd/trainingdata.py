"""Training data of the closure surrogate in the reference's closure formats.

A closure snapshot is what main_step04_liquid_closure.cpp:61-71 writes for
one call of run_liquid_closure, under a stem (the GPU benchmark driver uses
``<out>_closure_stepNNNN`` for each step of ``benchmark_closureDumpSteps``):

- ``_cellfields_ref.bin``, ``_cellfields_red.bin``: the reference and red
  states' CellFields (dump_io.cpp:19-39 layout);
- ``_pdrive_raw.bin``, ``_pdrive.bin``, ``_durepairx.bin``, ``_durepairy.bin``,
  ``_dueosx.bin``, ``_dueosy.bin``: f64 grids with an int32 Nx, Ny header
  (liquid_closure.cpp:446-462);
- ``_interzone_mask.bin``: the i32 interzone mask;
- ``_closure.kv``: ``step`` and ``betaRepairApplied`` (the applied beta,
  which no grid holds).

The readers return numpy arrays; ``training_set`` turns snapshots into the
surrogate's (features, target) pairs with the GPU kernels that sample
training data during a run (``gpu.closure_training_pair``), so a model
trained from files sees exactly what one trained in memory sees.
"""
from __future__ import annotations

import glob
import os
import re
from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import fileio

GRID_FILES = {"pdriveRaw": "_pdrive_raw.bin", "pdrive": "_pdrive.bin", "duRepairX": "_durepairx.bin",
              "duRepairY": "_durepairy.bin", "duEOSX": "_dueosx.bin", "duEOSY": "_dueosy.bin"}


def read_kv(path: str) -> Dict[str, str]:
    out = {}
    with open(path) as fh:
        for line in fh:
            line = line.strip()
            if line and "=" in line and not line.startswith("#"):
                k, v = line.split("=", 1)
                out[k.strip()] = v.strip()
    return out


def read_closure_snapshot(stem: str) -> Dict[str, object]:
    """One snapshot: {"ref", "red": CellFields dicts, the six grids,
    "interzoneMask", "betaRepairApplied", "step", "Nx", "Ny"}."""
    out: Dict[str, object] = {"ref": fileio.read_cell_fields(stem + "_cellfields_ref.bin"),
                              "red": fileio.read_cell_fields(stem + "_cellfields_red.bin")}
    shape = None
    for k, suf in GRID_FILES.items():
        Nx, Ny, g = fileio.read_grid(stem + suf, "<f8")
        if shape is None:
            shape = (Nx, Ny)
        elif shape != (Nx, Ny):
            raise RuntimeError(f"grid size mismatch in {stem + suf}")
        out[k] = g
    out["interzoneMask"] = fileio.read_grid(stem + "_interzone_mask.bin", "<i4")[2]
    kv = read_kv(stem + "_closure.kv")
    out["betaRepairApplied"] = float(kv["betaRepairApplied"])
    out["step"] = int(kv["step"])
    out["Nx"], out["Ny"] = shape
    return out


def list_closure_snapshots(prefix: str) -> List[str]:
    """Stems of the snapshots a driver run with output prefix ``prefix``
    wrote, in step order."""
    stems = []
    for kv in glob.glob(glob.escape(prefix) + "_closure_step*_closure.kv"):
        stem = kv[: -len("_closure.kv")]
        m = re.search(r"_closure_step(\d+)$", stem)
        if m:
            stems.append((int(m.group(1)), stem))
    return [s for _, s in sorted(stems)]


def training_set(p, stems: Sequence[str], out_scale: float) -> Tuple[np.ndarray, np.ndarray]:
    """features [S, Nc, 4] and targets [S, Nc, 2] (float32) of the snapshots,
    computed on the GPU."""
    from . import gpu

    feats, tgts = [], []
    for stem in stems:
        snap = read_closure_snapshot(stem)
        if (snap["Nx"], snap["Ny"]) != (p.Nx, p.Ny):
            raise ValueError(f"{stem}: grid {snap['Nx']}x{snap['Ny']} does not match the parameters")
        f, t = gpu.closure_training_pair(p, snap["ref"], snap, snap["betaRepairApplied"], out_scale)
        feats.append(f)
        tgts.append(t)
    if not feats:
        return np.zeros((0, p.Nc, 4), np.float32), np.zeros((0, p.Nc, 2), np.float32)
    return np.stack(feats), np.stack(tgts)


def exists(stem: str) -> bool:
    return os.path.exists(stem + "_closure.kv")
