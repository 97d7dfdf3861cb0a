"""The closure training-data readers on files written in the reference's
formats (CellFields: int32 Nx, Ny, N[Nc] int32, ten f64 fields; grids:
int32 Nx, Ny, Nc values), built here byte by byte with numpy."""
import os

import numpy as np
import pytest

from src_incompressible_openmp_amd import trainingdata
from src_incompressible_openmp_amd.fileio import CELL_FIELD_NAMES


def write_snapshot(stem, Nx, Ny, step, beta, rng):
    Nc = Nx * Ny
    fields = {}
    for part in ("ref", "red"):
        f = {"N": rng.integers(0, 20, Nc).astype("<i4")}
        for k in CELL_FIELD_NAMES:
            f[k] = rng.standard_normal(Nc)
        with open(f"{stem}_cellfields_{part}.bin", "wb") as fh:
            fh.write(np.array([Nx, Ny], "<i4").tobytes())
            fh.write(f["N"].tobytes())
            for k in CELL_FIELD_NAMES:
                fh.write(f[k].astype("<f8").tobytes())
        fields[part] = f
    grids = {}
    for k, suf in trainingdata.GRID_FILES.items():
        grids[k] = rng.standard_normal(Nc)
        with open(stem + suf, "wb") as fh:
            fh.write(np.array([Nx, Ny], "<i4").tobytes() + grids[k].astype("<f8").tobytes())
    mask = rng.integers(0, 2, Nc).astype("<i4")
    with open(stem + "_interzone_mask.bin", "wb") as fh:
        fh.write(np.array([Nx, Ny], "<i4").tobytes() + mask.tobytes())
    with open(stem + "_closure.kv", "w") as fh:
        fh.write(f"step={step}\nbetaRepairApplied={beta!r}\n")
    return fields, grids, mask


def test_read_closure_snapshot_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    stem = os.path.join(str(tmp_path), "o_closure_step0007")
    fields, grids, mask = write_snapshot(stem, 6, 4, 7, 0.123456789012345678, rng)
    s = trainingdata.read_closure_snapshot(stem)
    assert (s["Nx"], s["Ny"], s["step"]) == (6, 4, 7)
    assert s["betaRepairApplied"] == 0.123456789012345678
    for part in ("ref", "red"):
        assert np.array_equal(s[part]["N"], fields[part]["N"])
        for k in CELL_FIELD_NAMES:
            assert s[part][k].tobytes() == fields[part][k].tobytes()
    for k in trainingdata.GRID_FILES:
        assert s[k].tobytes() == grids[k].tobytes()
    assert np.array_equal(s["interzoneMask"], mask)


def test_list_closure_snapshots_in_step_order(tmp_path):
    rng = np.random.default_rng(4)
    d = str(tmp_path)
    for step in (12, 3, 100):
        write_snapshot(os.path.join(d, f"run_closure_step{step:04d}"), 3, 2, step, 0.5, rng)
    write_snapshot(os.path.join(d, "other_closure_step0001"), 3, 2, 1, 0.5, rng)
    stems = trainingdata.list_closure_snapshots(os.path.join(d, "run"))
    assert [os.path.basename(s) for s in stems] == ["run_closure_step0003", "run_closure_step0012",
                                                     "run_closure_step0100"]


def test_mismatched_grid_sizes_raise(tmp_path):
    rng = np.random.default_rng(5)
    stem = os.path.join(str(tmp_path), "x_closure_step0001")
    write_snapshot(stem, 4, 4, 1, 0.5, rng)
    with open(stem + "_dueosy.bin", "wb") as fh:
        fh.write(np.array([2, 8], "<i4").tobytes() + np.zeros(16).tobytes())
    with pytest.raises(RuntimeError):
        trainingdata.read_closure_snapshot(stem)
