"""Closure training data on the GPU: the snapshots a resident run keeps of
its plain closures (reference and red CellFields, the six grids, beta)
against the oracle's closure of the same reference/red states; the
training pair computed from a snapshot against the pair the run samples in
memory; and the GPU benchmark driver's closure dumps (main_step04's file
formats) against the snapshots of a resident run over the same steps.

Fields and grids are bit for bit (as in the closure's own parity test);
beta to 1e-12 relative (its sums are tree reductions on the GPU)."""
import os
import subprocess

import numpy as np
import pytest

from src_incompressible_openmp_amd.fileio import uniform_state, write_u8, write_xy
from src_incompressible_openmp_amd.params import to_struct, write_params_kv
from tests.golden.cases import case_params, case_state

pytestmark = pytest.mark.gpu

GRIDS = ("pdriveRaw", "pdrive", "duRepairX", "duRepairY", "duEOSX", "duEOSY")


@pytest.fixture(scope="module")
def gpu():
    from src_incompressible_openmp_amd import gpu as g

    g.load()
    assert g.device_count() >= 1, "no GPU visible"
    return g


def same(a, b) -> bool:
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def snapshots_equal(a, b):
    for part in ("ref", "red"):
        for k, arr in b[part].items():
            if isinstance(arr, np.ndarray):
                assert same(a[part][k], arr), (part, k)
    for k in GRIDS + ("interzoneMask",):
        assert same(a[k], b[k]), k
    assert a["betaRepairApplied"] == b["betaRepairApplied"]
    assert a["step"] == b["step"]


@pytest.mark.parametrize("case", ("small", "bcmix", "override"))
def test_snapshot_matches_oracle_closure(gpu, oracle, case):
    p = to_struct(case_params(case))
    x, v, _, _ = case_state(case)
    sim = gpu.Sim(p)
    try:
        sim.set_state(x, v)
        assert sim.closure_snapshot(True, fetch=True) is None  # none kept yet
        sim.step(1, 2)
        st = sim.step_stages(3)
        snap = sim.closure_snapshot(True, fetch=True)
    finally:
        sim.close()
    assert snap["step"] == 3
    o = oracle.closure(p, st["x_ref"], st["v_ref"], st["x_red"], st["v_red"])
    for part in ("ref", "red"):
        for k, arr in o["fields"][part].items():
            assert same(snap[part][k], arr), (case, part, k)
    for k in GRIDS:
        assert same(snap[k], o[k]), (case, k)
    assert np.array_equal(snap["interzoneMask"], o["mask"])
    b = o["metrics"]["betaRepairApplied"]
    assert abs(snap["betaRepairApplied"] - b) <= 1e-12 * max(abs(b), 1e-300)


def test_training_pair_matches_in_memory_sample(gpu):
    p = to_struct(case_params("small"))
    x, v, _, _ = case_state("small")
    sim = gpu.Sim(p)
    try:
        sim.set_state(x, v)
        sim.closure_snapshot(True)
        sim.closure_sample(True, 0.05)
        sim.step(1, 4)
        f_mem, t_mem = sim.closure_sample(True, 0.05, fetch=True)
        snap = sim.closure_snapshot(True, fetch=True)
    finally:
        sim.close()
    f, t = gpu.closure_training_pair(p, snap["ref"], snap, snap["betaRepairApplied"], 0.05)
    assert same(f, f_mem) and same(t, t_mem)


def test_driver_closure_dumps(gpu, tmp_path):
    """Six steps, metrics rows at 3 and 6, closure dumps at 2 and 5 (the
    steps between rows run batched)."""
    from src_incompressible_openmp_amd import build, trainingdata

    prm = case_params("small")
    p = to_struct(prm)
    d = str(tmp_path)
    x, v, t, r0 = uniform_state(p.n, p.Lx, p.Ly, seed=1)
    write_params_kv(os.path.join(d, "params.kv"), prm,
                    {"benchmark_nSteps": 6, "benchmark_metricsEvery": 3, "benchmark_dumpSteps": "-1",
                     "benchmark_closureDumpSteps": "2,5"})
    write_xy(os.path.join(d, "x.bin"), x, p.n)
    write_xy(os.path.join(d, "v.bin"), v, p.n)
    write_u8(os.path.join(d, "t.bin"), t, p.n)
    write_xy(os.path.join(d, "r0.bin"), r0, p.n)
    subprocess.run([build.DRIVER, os.path.join(d, "params.kv"), os.path.join(d, "x.bin"), os.path.join(d, "v.bin"),
                    os.path.join(d, "t.bin"), os.path.join(d, "r0.bin"), os.path.join(d, "o"), "1", "1", "1"],
                   check=True, timeout=60)
    stems = trainingdata.list_closure_snapshots(os.path.join(d, "o"))
    assert [os.path.basename(s) for s in stems] == ["o_closure_step0002", "o_closure_step0005"]
    sim = gpu.Sim(p)
    try:
        sim.set_state(x, v)
        sim.closure_snapshot(True)
        sim.step(1, 2)
        ref2 = sim.closure_snapshot(True, fetch=True)
        sim.step(3, 3)
        ref5 = sim.closure_snapshot(True, fetch=True)
    finally:
        sim.close()
    for stem, ref in zip(stems, (ref2, ref5)):
        snapshots_equal(trainingdata.read_closure_snapshot(stem), ref)
    f, tg = trainingdata.training_set(p, stems, 0.05)
    assert f.shape == (2, p.Nc, 4) and tg.shape == (2, p.Nc, 2)
    f5, t5 = gpu.closure_training_pair(p, ref5["ref"], ref5, ref5["betaRepairApplied"], 0.05)
    assert same(f[1], f5) and same(tg[1], t5)
