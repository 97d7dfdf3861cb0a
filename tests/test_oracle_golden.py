"""Pin the CPU oracle (libstdc++ draw mode) bitwise against the golden
fixtures written by the compiled original reference (tests/golden/)."""
import hashlib
import json
import os

import numpy as np
import pytest

from src_incompressible_openmp_amd.params import LAYOUT_BASE, LAYOUT_SHIFTED, to_struct
from tests.golden.cases import DIGEST_CASES, SMALL_CASES, case_params, case_state
from tests.oracle_lib import STD

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, f"{name}.npz"), allow_pickle=False)


def kv(g, prefix):
    return dict(zip(g[prefix + "_kv_keys"].tolist(), g[prefix + "_kv_vals"].tolist()))


def bits_equal(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def close6(a, b):
    # runout.kv values are printed with 6 significant digits
    return abs(a - b) <= 1e-5 * max(1.0, abs(b))


@pytest.fixture(params=SMALL_CASES)
def case(request):
    return request.param


def test_inputs_regenerate(case):
    g = load(case)
    x, v, _, _ = case_state(case)
    assert bits_equal(x, g["x0"]) and bits_equal(v, g["v0"])


def test_step01_stream_bc_collision(oracle, case):
    """main_step01_stream_bc_collision.cpp:34-55"""
    g = load(case)
    p = to_struct(case_params(case))
    x, v = g["x0"].copy(), g["v0"].copy()
    oracle.force_stream(p, x, v)
    wi = oracle.apply_bc(STD, p, x, v, 104729)
    cid, _ = oracle.shift_cell_ids(STD, p, x, 130363)
    oracle.collide(STD, p, v, cid, 433494437)
    assert bits_equal(x, g["s1_x"])
    assert bits_equal(v, g["s1_v"])
    cf = oracle.cell_fields(p, x, v, p.gamma / ((p.Lx / p.Nx) * (p.Ly / p.Ny)))
    for k in cf:
        assert bits_equal(cf[k], g["s1_cf_" + k]), k
    r = kv(g, "s1")
    denom = max(p.Lx * p.dt, 1e-30)
    assert close6(wi["dPyBot"] / denom, r["pBotInst"])
    assert close6(wi["dPyTop"] / denom, r["pTopInst"])


@pytest.mark.parametrize("layout", ["base", "shifted"])
def test_step02_gauss_seidel_pass(oracle, case, layout):
    """main_step02_zone_pass.cpp:33 -> run_zone_pass (zone_pass.cpp:1096-1163)"""
    g = load(case)
    p = to_struct(case_params(case))
    x, v = g["x0"].copy(), g["v0"].copy()
    lay = LAYOUT_BASE if layout == "base" else LAYOUT_SHIFTED
    m, owner = oracle.zone_pass(STD, p, x, v, lay, 88172645463325252, gauss_seidel=True)
    assert bits_equal(x, g[f"s2{layout}_x"])
    assert bits_equal(v, g[f"s2{layout}_v"])
    assert np.array_equal(owner, g[f"s2{layout}_owner"])
    r = kv(g, f"s2{layout}")
    assert m["nParticlesMovedDense"] == r["nParticlesMovedDense"]
    assert m["nParticlesMovedSparse"] == r["nParticlesMovedSparse"]
    assert m["nZones"] == r["nZonesExecuted"]
    assert close6(m["corrRmsDU"], r["corrRmsDU"]) and close6(m["corrMaxDU"], r["corrMaxDU"])


def test_step04_closure(oracle, case):
    """main_step04_liquid_closure.cpp:54 -> run_liquid_closure"""
    g = load(case)
    p = to_struct(case_params(case))
    out = oracle.closure(p, g["x0"], g["v0"], g["s4_xred"], g["s4_vred"])
    assert bits_equal(out["x"], g["s4_x"])
    assert bits_equal(out["v"], g["s4_v"])
    names = {"pdriveRaw": "pdrive_raw", "pdrive": "pdrive", "duRepairX": "durepairx", "duRepairY": "durepairy",
             "duEOSX": "dueosx", "duEOSY": "dueosy"}
    for ours, theirs in names.items():
        assert bits_equal(out[ours], g["s4_" + theirs]), ours
    assert np.array_equal(out["mask"], g["s4_mask"])
    for f in ("ref", "red", "repaired", "out"):
        for k, a in out["fields"][f].items():
            assert bits_equal(a, g[f"s4_cf{f}_{k}"]), (f, k)
    r = kv(g, "s4")
    m = out["metrics"]
    for k in ("betaRepairOpt", "betaRepairApplied", "meanPdriveRaw", "meanPdrive", "rmsRepairDU", "maxRepairDU",
              "rmsEOSDU", "maxEOSDU", "momentumDeltaX", "momentumDeltaY"):
        assert close6(m[k], r[k]), k
    assert m["nMaskedInterzoneCells"] == r["nMaskedInterzoneCells"]


def test_step05_full_step_gauss_seidel(oracle, case):
    """main_step05_full_step.cpp:84-111 (Gauss-Seidel passes, step seeds 0)"""
    g = load(case)
    p = to_struct(case_params(case))
    x, v = g["x0"].copy(), g["v0"].copy()
    oracle.force_stream(p, x, v)
    oracle.apply_bc(STD, p, x, v, 104729)
    cid, _ = oracle.shift_cell_ids(STD, p, x, 130363)
    oracle.collide(STD, p, v, cid, 433494437)
    assert bits_equal(x, g["s5_ref_x"]) and bits_equal(v, g["s5_ref_v"])
    xr, vr = x.copy(), v.copy()
    oracle.zone_pass(STD, p, x, v, LAYOUT_BASE, 88172645463325252, gauss_seidel=True)
    assert bits_equal(x, g["s5_after_base_x"]) and bits_equal(v, g["s5_after_base_v"])
    oracle.zone_pass(STD, p, x, v, LAYOUT_SHIFTED, 88172645463325252 + 4099, gauss_seidel=True)
    assert bits_equal(x, g["s5_red_x"]) and bits_equal(v, g["s5_red_v"])
    out = oracle.closure(p, xr, vr, x, v)
    assert bits_equal(out["x"], g["s5_x"]) and bits_equal(out["v"], g["s5_v"])


def test_step09_snapshot_step(oracle, case):
    """main_step09_openmp_full_validate_syncserial.cpp:61-67: snapshot passes;
    the serial and OpenMP outputs of the original are identical."""
    g = load(case)
    p = to_struct(case_params(case))
    x, v = g["x0"].copy(), g["v0"].copy()
    st = oracle.step(STD, p, x, v, 0, stages=True)
    assert bits_equal(st["x_ref"], g["s9_ref_x"]) and bits_equal(st["v_ref"], g["s9_ref_v"])
    assert bits_equal(st["x_base"], g["s9_serial_after_base_x"])
    assert bits_equal(st["v_base"], g["s9_serial_after_base_v"])
    assert bits_equal(x, g["s9_serial_x"]) and bits_equal(v, g["s9_serial_v"])
    assert bits_equal(g["s9_omp_x"], g["s9_serial_x"]) and bits_equal(g["s9_omp_v"], g["s9_serial_v"])
    r = kv(g, "s9")
    assert st["base"]["nParticlesMovedDense"] == r["serial_base_nParticlesMovedDense"]
    assert st["shifted"]["nParticlesMovedSparse"] == r["serial_shifted_nParticlesMovedSparse"]


def test_benchmark_driver_steps(oracle, case):
    """main_benchmark_poiseuille_openmp_profiled_shifted_profiled.cpp:237-279,
    steps 1..6, final dump and per-step metrics."""
    g = load(case)
    p = to_struct(case_params(case))
    x, v = g["x0"].copy(), g["v0"].copy()
    hdr = g["bench_metrics_header"].tolist()
    rows = g["bench_metrics"]
    for step in range(1, int(g["bench_steps"]) + 1):
        st = oracle.step(STD, p, x, v, step)
        row = dict(zip(hdr, rows[step - 1]))
        moved_d = st["base"]["nParticlesMovedDense"] + st["shifted"]["nParticlesMovedDense"]
        moved_s = st["base"]["nParticlesMovedSparse"] + st["shifted"]["nParticlesMovedSparse"]
        assert moved_d == row["particlesMovedDense"] and moved_s == row["particlesMovedSparse"]
        assert close6(st["closure"]["betaRepairApplied"], row["betaRepairApplied"])
        ke = float(np.mean(0.5 * (v[0::2] ** 2 + v[1::2] ** 2)))
        assert close6(ke, row["meanKinetic"])
    assert bits_equal(x, g["bench_x"]) and bits_equal(v, g["bench_v"])


@pytest.mark.parametrize("name", DIGEST_CASES)
def test_digest_case(oracle, name):
    """configs-scale case (50x50 cells, 25,000 particles): SHA-256 of the
    original's outputs."""
    with open(os.path.join(GOLD, "digests.json")) as fh:
        d = json.load(fh)[name]
    p = to_struct(case_params(name))
    x, v, _, _ = case_state(name)

    def h(a):
        return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()

    assert h(x) == d["x0"] and h(v) == d["v0"]
    x9, v9 = x.copy(), v.copy()
    st = oracle.step(STD, p, x9, v9, 0, stages=True)
    assert h(st["x_ref"]) == d["s9_ref_x"] and h(st["v_ref"]) == d["s9_ref_v"]
    assert h(st["x_base"]) == d["s9_serial_after_base_x"] and h(st["v_base"]) == d["s9_serial_after_base_v"]
    assert h(x9) == d["s9_serial_x"] and h(v9) == d["s9_serial_v"]
    for step in range(1, d["bench_steps"] + 1):
        oracle.step(STD, p, x, v, step)
    assert h(x) == d["bench_x"] and h(v) == d["bench_v"]
