"""Parity at the benchmark's scale: the 1024 x 512-cell channel with 20
particles per cell (10,485,760 particles, 861 base and 882 shifted zones),
steps 1-5 and every 100th step of a 1000-step run, stage by stage against
the oracle's tile-local step (counter mode).  Each checked step starts both
codes from the same state (replay), so every stage is compared at its own
precision: positions, the states after the collision and after each zone
pass, the moved-particle counts and the owner maps bit for bit; the
velocities after the closure (beta is a tree sum on the GPU) to 1e-12
relative."""
import time

import numpy as np
import pytest

from src_incompressible_openmp_amd.fileio import uniform_state
from src_incompressible_openmp_amd.params import LAYOUT_BASE, LAYOUT_SHIFTED, poiseuille_params, to_struct
from tests.oracle_lib import CTR

pytestmark = pytest.mark.gpu


@pytest.mark.timeout(900)
def test_scale_1024x512_steps_stage_by_stage(oracle):
    from src_incompressible_openmp_amd import gpu

    gpu.load()
    p = to_struct(poiseuille_params(Nx=1024, Ny=512, gamma=20.0))
    for layout in (LAYOUT_BASE, LAYOUT_SHIFTED):
        og, _ = gpu.build_zone_layout(p, layout)
        oo, nz = oracle.layout(p, layout)
        assert np.array_equal(og, oo) and nz == (861 if layout == LAYOUT_BASE else 882)
    x, v, _, _ = uniform_state(p.n, p.Lx, p.Ly, seed=1)
    sim = gpu.Sim(p)
    try:
        for s in (1, 2, 3, 4, 5):
            print(f"scale step {s}: n={p.n}", flush=True)
            sim.set_state(x, v)
            replay_step(sim, oracle, p, x, v, s)
    finally:
        sim.close()


def replay_step(sim, oracle, p, x, v, s):
    """Step s from state (x, v) on the GPU (stages) and in the oracle, which
    advances x, v in place; every stage compared."""
    sg = sim.step_stages(s)
    xg, vg = sim.get_state()
    t0 = time.perf_counter()
    so = oracle.step(CTR, p, x, v, s, stages=True)
    print(f"  oracle step {time.perf_counter() - t0:.1f} s, moved", so["base"]["nParticlesMovedDense"],
          so["shifted"]["nParticlesMovedDense"], flush=True)
    for a in ("ref", "base", "red"):
        assert sg["x_" + a].tobytes() == so["x_" + a].tobytes(), (s, a)
        assert sg["v_" + a].tobytes() == so["v_" + a].tobytes(), (s, a)
    assert xg.tobytes() == x.tobytes(), s
    assert float(np.max(np.abs(vg - v))) <= 1e-12 * max(1.0, float(np.max(np.abs(v)))), s
    m = sg["metrics"]
    for L in ("base", "shifted"):
        for k in ("nZones", "nParticlesMovedDense", "nParticlesMovedSparse"):
            assert m[L][k] == so[L][k], (s, L, k)
        assert m[L]["corrRmsDU"] == so[L]["corrRmsDU"] and m[L]["corrMaxDU"] == so[L]["corrMaxDU"], (s, L)
    assert m["wall"]["nBot"] == so["wall"]["nBot"] and m["wall"]["nTop"] == so["wall"]["nTop"], s
    for k in ("betaRepairApplied", "betaRepairNum", "betaRepairDen"):
        assert abs(m["closure"][k] - so["closure"][k]) <= 1e-10 * max(1.0, abs(so["closure"][k])), (s, k)


@pytest.mark.timeout(900)
def test_scale_1024x512_every_100th_step_of_1000(oracle):
    """A 1000-step resident run (the benchmark's fused step form); at steps
    100, 200, ..., 1000 the step is replayed from the run's state on the GPU
    stage by stage and in the oracle.  The run then continues from the GPU's
    own result, so the checked states are those of the developing flow."""
    from src_incompressible_openmp_amd import gpu

    gpu.load()
    p = to_struct(poiseuille_params(Nx=1024, Ny=512, gamma=20.0))
    x, v, _, _ = uniform_state(p.n, p.Lx, p.Ly, seed=2)
    sim = gpu.Sim(p)
    try:
        sim.set_state(x, v)
        done = 0
        for s in range(100, 1001, 100):
            sim.step(done + 1, s - 1 - done)
            xs, vs = sim.get_state()
            print(f"scale run: step {s} replayed from the run's state", flush=True)
            sim.set_state(xs, vs)
            replay_step(sim, oracle, p, xs, vs, s)
            done = s
    finally:
        sim.close()


def bits_equal(a: np.ndarray, b: np.ndarray) -> bool:
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


@pytest.mark.timeout(900)
def test_scale_4096x2048_first_step_stage_by_stage(oracle):
    """The large single instance: 4096 x 2048 cells x 20 particles per cell
    (167,772,160 particles, 13,448 base zones), its first step stage by
    stage against the oracle, at the same precision as above."""
    from src_incompressible_openmp_amd import gpu

    gpu.load()
    p = to_struct(poiseuille_params(Nx=4096, Ny=2048, gamma=20.0))
    og, nzg = gpu.build_zone_layout(p, LAYOUT_BASE)
    oo, nz = oracle.layout(p, LAYOUT_BASE)
    assert np.array_equal(og, oo) and len(nzg) == nz
    print(f"large step 1: n={p.n}, {nz} base zones", flush=True)
    x, v, _, _ = uniform_state(p.n, p.Lx, p.Ly, seed=1)
    sim = gpu.Sim(p)
    try:
        sim.set_state(x, v)
        sg = sim.step_stages(1)
        xg, vg = sim.get_state()
    finally:
        sim.close()
    print("  GPU step done", flush=True)
    t0 = time.perf_counter()
    so = oracle.step(CTR, p, x, v, 1, stages=True)
    print(f"  oracle step {time.perf_counter() - t0:.1f} s, moved", so["base"]["nParticlesMovedDense"],
          so["shifted"]["nParticlesMovedDense"], flush=True)
    for a in ("ref", "base", "red"):
        assert bits_equal(sg["x_" + a], so["x_" + a]), a
        assert bits_equal(sg["v_" + a], so["v_" + a]), a
    assert bits_equal(xg, x)
    assert float(np.max(np.abs(vg - v))) <= 1e-12 * max(1.0, float(np.max(np.abs(v))))
    m = sg["metrics"]
    for L in ("base", "shifted"):
        for k in ("nZones", "nParticlesMovedDense", "nParticlesMovedSparse"):
            assert m[L][k] == so[L][k], (L, k)
    for k in ("betaRepairApplied", "betaRepairNum", "betaRepairDen"):
        assert abs(m["closure"][k] - so["closure"][k]) <= 1e-10 * max(1.0, abs(so["closure"][k])), k
