"""Drop-in check: the reference's own drivers, relinked against this
library's mpcd:: API (oracle/dropin.mk), run on the GPU and agree with
the oracle's counter mode on the same inputs.  Same bar as the C-ABI parity
tests: positions bitwise, velocities through the closure's reduced beta to
1e-12 (1e-9 after several steps)."""
import os
import subprocess

import numpy as np
import pytest

from src_incompressible_openmp_amd import fileio
from src_incompressible_openmp_amd.params import to_struct, write_params_kv
from tests.golden.cases import case_params, case_state, ref_red_pair
from tests.oracle_lib import CTR

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "dropin")


def exe(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not built (oracle/dropin.mk needs the reference sources)")
    return path


def bits(a, b):
    return np.ascontiguousarray(a).tobytes() == np.ascontiguousarray(b).tobytes()


def vclose(a, b, rtol):
    return float(np.max(np.abs(a - b))) <= rtol * max(1.0, float(np.max(np.abs(b))))


def inputs(d, case, extra=None):
    prm = case_params(case)
    x, v, t, r0 = case_state(case)
    n = int(prm["n"])
    write_params_kv(os.path.join(d, "params.kv"), prm, extra)
    fileio.write_xy(os.path.join(d, "x.bin"), x, n)
    fileio.write_xy(os.path.join(d, "v.bin"), v, n)
    fileio.write_u8(os.path.join(d, "type.bin"), t, n)
    fileio.write_xy(os.path.join(d, "r0.bin"), r0, n)
    return prm, x, v, [os.path.join(d, f) for f in ("params.kv", "x.bin", "v.bin", "type.bin", "r0.bin")]


def run(path, *args):
    print("dropin:", os.path.basename(path), flush=True)
    subprocess.run([path, *args], check=True, timeout=120, stdout=subprocess.DEVNULL,
                   env=dict(os.environ, OMP_NUM_THREADS="1"))


@pytest.mark.parametrize("case", ["small", "bcmix"])
def test_step01_stream_bc_collision(oracle, tmp_path, case):
    prm, x, v, ins = inputs(str(tmp_path), case)
    o = str(tmp_path / "o")
    run(exe("main_step01_stream_bc_collision"), *ins, o, "1", "1")
    p = to_struct(prm)
    oracle.force_stream(p, x, v)
    oracle.apply_bc(CTR, p, x, v, 104729)
    cid, _ = oracle.shift_cell_ids(CTR, p, x, 130363)
    oracle.collide(CTR, p, v, cid, 433494437)
    assert bits(fileio.read_xy(o + "_x.bin", p.n), x)
    vd = fileio.read_xy(o + "_v.bin", p.n)
    # the mean-flow removal sums velocities in a tree: 1e-12, else bitwise
    assert vclose(vd, v, 1e-12) if p.keepMeanFlow else bits(vd, v)


@pytest.mark.parametrize("case", ["small", "override"])
def test_step04_liquid_closure(oracle, tmp_path, case):
    prm, x, v, ins = inputs(str(tmp_path), case)
    p = to_struct(prm)
    xr, vr = ref_red_pair(x, v, p.Lx, p.Ly)
    fileio.write_xy(str(tmp_path / "xr.bin"), xr, p.n)
    fileio.write_xy(str(tmp_path / "vr.bin"), vr, p.n)
    o = str(tmp_path / "o")
    run(exe("main_step04_liquid_closure"), ins[0], ins[1], ins[2], str(tmp_path / "xr.bin"),
        str(tmp_path / "vr.bin"), ins[3], ins[4], o, "1", "1", "1")
    ref = oracle.closure(p, x, v, xr, vr)
    assert bits(fileio.read_xy(o + "_x.bin", p.n), ref["x"])
    assert vclose(fileio.read_xy(o + "_v.bin", p.n), ref["v"], 1e-12)
    for f, name in (("pdrive_raw", "pdriveRaw"), ("pdrive", "pdrive"), ("durepairx", "duRepairX"),
                    ("durepairy", "duRepairY"), ("dueosx", "duEOSX"), ("dueosy", "duEOSY")):
        assert bits(fileio.read_grid(o + f"_{f}.bin", "<f8")[2], ref[name]), f
    assert np.array_equal(fileio.read_grid(o + "_interzone_mask.bin", "<i4")[2], ref["mask"])
    for f in ("ref", "red"):
        got = fileio.read_cell_fields(o + f"_cellfields_{f}.bin")
        for k, a in ref["fields"][f].items():
            assert bits(got[k], a), (f, k)
    for f in ("repaired", "out"):
        got = fileio.read_cell_fields(o + f"_cellfields_{f}.bin")
        assert np.array_equal(got["N"], ref["fields"][f]["N"])
        for k in ("Ux", "Uy", "Px", "Py"):
            assert vclose(got[k], ref["fields"][f][k], 1e-12), (f, k)


@pytest.mark.parametrize("case", ["small", "single"])
def test_step09_snapshot_step(oracle, tmp_path, case):
    """ref build, snapshot base + shifted passes (serial and "OpenMP" entry
    points) and the closure: the driver's step equals the oracle's step 0."""
    prm, x, v, ins = inputs(str(tmp_path), case)
    o = str(tmp_path / "o")
    run(exe("main_step09_openmp_full_validate_syncserial"), *ins, o, "1", "1", "2")
    p = to_struct(prm)
    st = oracle.step(CTR, p, x, v, 0, stages=True)
    assert bits(fileio.read_xy(o + "_ref_x.bin", p.n), st["x_ref"])
    assert bits(fileio.read_xy(o + "_ref_v.bin", p.n), st["v_ref"])
    for tag in ("serial_after_base", "omp_after_base"):
        assert bits(fileio.read_xy(o + f"_{tag}_x.bin", p.n), st["x_base"]), tag
        assert bits(fileio.read_xy(o + f"_{tag}_v.bin", p.n), st["v_base"]), tag
    for tag in ("serial", "omp"):
        assert bits(fileio.read_xy(o + f"_{tag}_x.bin", p.n), x), tag
        assert vclose(fileio.read_xy(o + f"_{tag}_v.bin", p.n), v, 1e-12), tag


def test_benchmark_driver_six_steps(oracle, tmp_path):
    """The production benchmark driver, steps 1..6 through the drop-in API."""
    steps = 6
    prm, x, v, ins = inputs(str(tmp_path), "small", {"benchmark_nSteps": steps, "benchmark_metricsEvery": 1,
                                                     "benchmark_dumpSteps": str(steps)})
    o = str(tmp_path / "o")
    run(exe("main_benchmark_poiseuille_openmp_profiled_shifted_profiled"), *ins, o, "1", "1", "2")
    p = to_struct(prm)
    for s in range(1, steps + 1):
        oracle.step(CTR, p, x, v, s)
    stem = o + f"_step{steps:04d}"
    assert np.max(np.abs(fileio.read_xy(stem + "_x.bin", p.n) - x)) < 1e-9
    assert vclose(fileio.read_xy(stem + "_v.bin", p.n), v, 1e-9)


@pytest.mark.parametrize("layout", ["base", "shifted"])
def test_step02_gauss_seidel_pass(oracle, tmp_path, layout):
    """run_zone_pass through the drop-in: state, owner map and runout."""
    from src_incompressible_openmp_amd import gpu
    from src_incompressible_openmp_amd.params import LAYOUT_BASE, LAYOUT_SHIFTED
    from tests.test_occupancy_stats import counts

    prm, x, v, ins = inputs(str(tmp_path), "small")
    o = str(tmp_path / "o")
    run(exe("main_step02_zone_pass"), *ins, o, layout, "1", "1")
    p = to_struct(prm)
    x0 = x.copy()
    m, owner = oracle.zone_pass(CTR, p, x, v, LAYOUT_BASE if layout == "base" else LAYOUT_SHIFTED,
                                88172645463325252, gauss_seidel=True)
    assert bits(fileio.read_xy(o + "_x.bin", p.n), x)
    assert bits(fileio.read_xy(o + "_v.bin", p.n), v)
    assert np.array_equal(fileio.read_grid(o + "_zone_owner.bin", "<i4")[2], owner)
    kv = {}
    with open(o + "_runout.kv") as fh:
        for line in fh:
            if "=" in line:
                k, val = line.strip().split("=", 1)
                kv[k] = val
    assert int(float(kv["nParticlesMovedDense"])) == m["nParticlesMovedDense"]
    assert int(float(kv["nParticlesMovedSparse"])) == m["nParticlesMovedSparse"]
    for tag, xs in (("Before", x0), ("After", x)):
        s = gpu.occupancy_stats(p, counts(p, xs))
        assert abs(float(kv["occStd" + tag]) - s["occStd"]) <= 1e-5 * max(1.0, s["occStd"])
        assert int(float(kv["nDense" + tag])) == s["nDense"] and int(float(kv["nSparse" + tag])) == s["nSparse"]
