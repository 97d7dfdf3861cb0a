"""GPU parity: libmpcd_gpu.so (through its C ABI) against the oracle's
counter-based mode on the same inputs and the same seeds.

Bar: bitwise for positions, cell indices, occupancy and redistribution
moves, and for every FP64 quantity whose sum order the GPU reproduces (per
cell sums run over particles in ascending id, as in the reference); 1e-12
relative for velocities that pass through a tree reduction (the closure's
beta = num/den over all cells, the optional mean-flow removal).
"""
import numpy as np
import pytest

from src_incompressible_openmp_amd.params import LAYOUT_BASE, LAYOUT_SHIFTED, poiseuille_params, to_struct
from tests.golden.cases import SMALL_CASES, case_params, case_state, ref_red_pair
from tests.oracle_lib import CTR

pytestmark = pytest.mark.gpu

CASES = SMALL_CASES + ("cfg0",)
RTOL_V = 1e-12  # velocities after a tree-reduced scalar


@pytest.fixture(scope="module")
def gpu():
    from src_incompressible_openmp_amd import gpu as g

    g.load()
    assert g.device_count() >= 1, "no GPU visible"
    return g


def bits(a, b):
    return np.ascontiguousarray(a).tobytes() == np.ascontiguousarray(b).tobytes()


def vclose(a, b, rtol=RTOL_V):
    scale = max(1.0, float(np.max(np.abs(b))))
    return float(np.max(np.abs(a - b))) <= rtol * scale


def post_collision(oracle, case):
    p = to_struct(case_params(case))
    x, v, _, _ = case_state(case)
    oracle.force_stream(p, x, v)
    oracle.apply_bc(CTR, p, x, v, 104729)
    cid, _ = oracle.shift_cell_ids(CTR, p, x, 130363)
    oracle.collide(CTR, p, v, cid, 433494437)
    return p, x, v


@pytest.mark.parametrize("case", CASES)
def test_force_stream_bc(gpu, oracle, case):
    p = to_struct(case_params(case))
    x, v, _, _ = case_state(case)
    xo, vo = x.copy(), v.copy()
    oracle.force_stream(p, xo, vo)
    wo = oracle.apply_bc(CTR, p, xo, vo, 104729)
    xg, vg = x.copy(), v.copy()
    wg = gpu.force_stream_bc(p, xg, vg, 104729)
    assert bits(xg, xo) and bits(vg, vo)
    for k in ("nBot", "nTop", "nLeft", "nRight"):
        assert wg[k] == wo[k]
    for k in ("dPyBot", "dPyTop", "dEwall"):
        assert abs(wg[k] - wo[k]) <= 1e-9 * max(1.0, abs(wo[k]))


@pytest.mark.parametrize("case", CASES)
def test_shift_cell_ids(gpu, oracle, case):
    p = to_struct(case_params(case))
    x, _, _, _ = case_state(case)
    co, _ = oracle.shift_cell_ids(CTR, p, x, 130363)
    cg = gpu.srd_cell_id_with_random_shift(p, x, 130363)
    assert np.array_equal(cg, co)


@pytest.mark.parametrize("case", CASES)
def test_collision(gpu, oracle, case):
    p = to_struct(case_params(case))
    x, v, _, _ = case_state(case)
    cid, _ = oracle.shift_cell_ids(CTR, p, x, 130363)
    vo, vg = v.copy(), v.copy()
    oracle.collide(CTR, p, vo, cid, 433494437)
    gpu.srd_collision_step(p, vg, cid, 433494437)
    if p.keepMeanFlow:
        assert vclose(vg, vo)
    else:
        assert bits(vg, vo)


@pytest.mark.parametrize("case", CASES)
def test_cell_fields(gpu, oracle, case):
    p, x, v = post_collision(oracle, case)
    rt = p.gamma / ((p.Lx / p.Nx) * (p.Ly / p.Ny))
    fo = oracle.cell_fields(p, x, v, rt)
    fg = gpu.compute_cell_fields(p, x, v, rt)
    for k in fo:
        assert bits(fg[k], fo[k]), k


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("layout", [LAYOUT_BASE, LAYOUT_SHIFTED])
def test_zone_pass(gpu, oracle, case, layout):
    p, x, v = post_collision(oracle, case)
    xo, vo = x.copy(), v.copy()
    mo, owo = oracle.zone_pass(CTR, p, xo, vo, layout, 88172645463325252)
    xg, vg = x.copy(), v.copy()
    mg, owg = gpu.zone_pass(p, xg, vg, layout, 88172645463325252)
    assert np.array_equal(owg, owo)
    assert mg["nParticlesMovedDense"] == mo["nParticlesMovedDense"]
    assert mg["nParticlesMovedSparse"] == mo["nParticlesMovedSparse"]
    assert bits(xg, xo)
    assert bits(vg, vo)
    assert mg["corrRmsDU"] == mo["corrRmsDU"] and mg["corrMaxDU"] == mo["corrMaxDU"]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("layout", [LAYOUT_BASE, LAYOUT_SHIFTED])
def test_zone_pass_gauss_seidel(gpu, oracle, case, layout):
    """run_zone_pass: zones applied one after another to the evolving state."""
    p, x, v = post_collision(oracle, case)
    xo, vo = x.copy(), v.copy()
    mo, owo = oracle.zone_pass(CTR, p, xo, vo, layout, 88172645463325252, gauss_seidel=True)
    xg, vg = x.copy(), v.copy()
    mg, owg = gpu.zone_pass(p, xg, vg, layout, 88172645463325252, gauss_seidel=True)
    assert np.array_equal(owg, owo)
    assert mg["nParticlesMovedDense"] == mo["nParticlesMovedDense"]
    assert mg["nParticlesMovedSparse"] == mo["nParticlesMovedSparse"]
    assert bits(xg, xo) and bits(vg, vo)
    assert mg["corrRmsDU"] == mo["corrRmsDU"] and mg["corrMaxDU"] == mo["corrMaxDU"]


@pytest.mark.parametrize("case", ["small", "bcmix", "override"])
def test_single_zone_kernel(gpu, oracle, case):
    """run_zone_kernel_inplace on individual zones of both layouts."""
    p, x, v = post_collision(oracle, case)
    for layout in (LAYOUT_BASE, LAYOUT_SHIFTED):
        _, rects = gpu.build_zone_layout(p, layout)
        for zi in range(0, len(rects), max(1, len(rects) // 4)):
            seed = 9001 + zi
            xo, vo = x.copy(), v.copy()
            st, _, mv = oracle.zone_kernel(CTR, p, xo, vo, layout, zi, seed)
            xg, vg = x.copy(), v.copy()
            sg, mg = gpu.run_zone_kernel(p, xg, vg, rects[zi], seed)
            assert np.array_equal(mg, mv)
            assert bits(xg, xo) and bits(vg, vo)
            assert bits(sg, st)


@pytest.mark.parametrize("case", CASES)
def test_closure(gpu, oracle, case):
    p = to_struct(case_params(case))
    x, v, _, _ = case_state(case)
    xr, vr = ref_red_pair(x, v, p.Lx, p.Ly)
    o = oracle.closure(p, x, v, xr, vr)
    g = gpu.liquid_closure(p, x, v, xr, vr)
    for k in ("pdriveRaw", "pdrive", "duRepairX", "duRepairY", "duEOSX", "duEOSY"):
        assert bits(g[k], o[k]), k
    assert np.array_equal(g["mask"], o["mask"])
    assert bits(g["x"], o["x"])
    assert vclose(g["v"], o["v"])
    assert vclose(g["v_repaired"], o["v_repaired"])  # stateRepaired: beta is a tree-reduced ratio
    mo, mg = o["metrics"], g["metrics"]
    assert mg["nMaskedInterzoneCells"] == mo["nMaskedInterzoneCells"]
    for k in ("betaRepairNum", "betaRepairDen", "betaRepairOpt", "betaRepairApplied", "meanPdriveRaw", "meanPdrive",
              "rmsRepairDU", "maxRepairDU", "meanAbsRepairDU", "rmsEOSDU", "maxEOSDU"):
        assert abs(mg[k] - mo[k]) <= 1e-10 * max(1.0, abs(mo[k])), k


@pytest.mark.parametrize("case", CASES)
def test_full_step_stages(gpu, oracle, case):
    """One resident step (benchmark driver seeds, step 3) stage by stage."""
    p = to_struct(case_params(case))
    x, v, _, _ = case_state(case)
    xo, vo = x.copy(), v.copy()
    so = oracle.step(CTR, p, xo, vo, 3, stages=True)
    sim = gpu.Sim(p)
    sim.set_state(x, v)
    sg = sim.step_stages(3)
    xg, vg = sim.get_state()
    exact_v = not p.keepMeanFlow
    for a in ("ref", "base", "red"):
        assert bits(sg["x_" + a], so["x_" + a]), a
        if exact_v:
            assert bits(sg["v_" + a], so["v_" + a]), a
        else:
            assert vclose(sg["v_" + a], so["v_" + a]), a
    assert bits(xg, xo)
    assert vclose(vg, vo)
    m = sg["metrics"]
    for L in ("base", "shifted"):
        assert m[L]["nParticlesMovedDense"] == so[L]["nParticlesMovedDense"]
        assert m[L]["nParticlesMovedSparse"] == so[L]["nParticlesMovedSparse"]
    assert m["wall"]["nBot"] == so["wall"]["nBot"] and m["wall"]["nTop"] == so["wall"]["nTop"]


@pytest.mark.parametrize("case", ["small", "cfg0"])
def test_steps_replayed(gpu, oracle, case):
    """Steps 1..6: each step starts both codes from the oracle's state."""
    p = to_struct(case_params(case))
    x, v, _, _ = case_state(case)
    sim = gpu.Sim(p)
    for s in range(1, 7):
        sim.set_state(x, v)
        sim.step(s)
        xg, vg = sim.get_state()
        oracle.step(CTR, p, x, v, s)
        assert bits(xg, x), s
        assert vclose(vg, v), s


def test_resident_run_matches_replay(gpu, oracle):
    """Ten resident steps on the GPU equal ten oracle steps (the tiny beta
    rounding difference does not flip any cell decision on this case)."""
    p = to_struct(case_params("small"))
    x, v, _, _ = case_state("small")
    sim = gpu.Sim(p)
    sim.set_state(x, v)
    sim.step(1, 10)
    xg, vg = sim.get_state()
    for s in range(1, 11):
        oracle.step(CTR, p, x, v, s)
    assert np.max(np.abs(xg - x)) < 1e-9
    assert vclose(vg, v, 1e-9)


def test_determinism(gpu):
    p = to_struct(case_params("cfg0"))
    x, v, _, _ = case_state("cfg0")
    outs = []
    for _ in range(2):
        sim = gpu.Sim(p)
        sim.set_state(x, v)
        sim.step(1, 3)
        outs.append(sim.get_state())
    assert bits(outs[0][0], outs[1][0]) and bits(outs[0][1], outs[1][1])


def test_unsupported_wall_mode_raises(gpu):
    prm = case_params("small")
    prm["boundary_bottom"] = "sticky"
    p = to_struct(prm)
    x, v, _, _ = case_state("small")
    x = x.copy()
    x[1::2] -= 0.1  # some particles below the bottom wall
    with pytest.raises(gpu.MpcdError) as e:
        gpu.apply_bc_general(p, x, v.copy(), 1)
    assert e.value.code == -3
    # a mode nobody crosses is not an error (the reference throws only on a hit)
    prm["boundary_bottom"] = "thermalize"
    prm["boundary_top"] = "sticky"
    p = to_struct(prm)
    x2, v2, _, _ = case_state("small")
    gpu.apply_bc_general(p, x2, v2, 1)


def test_disallowed_zone_modes_raise(gpu):
    prm = case_params("small")
    prm["useZoneRedistribution"] = False
    p = to_struct(prm)
    x, v, _, _ = case_state("small")
    with pytest.raises(gpu.MpcdError):
        gpu.zone_pass(p, x, v, LAYOUT_BASE, 1)


def test_medium_grid_step(gpu, oracle):
    """256x128 cells, 20 per cell (655,360 particles), one full step."""
    prm = poiseuille_params(Nx=256, Ny=128, gamma=20.0)
    p = to_struct(prm)
    from src_incompressible_openmp_amd.fileio import uniform_state

    x, v, _, _ = uniform_state(p.n, p.Lx, p.Ly, seed=11)
    xo, vo = x.copy(), v.copy()
    so = oracle.step(CTR, p, xo, vo, 1, stages=True)
    sim = gpu.Sim(p)
    sim.set_state(x, v)
    sg = sim.step_stages(1)
    xg, vg = sim.get_state()
    for a in ("ref", "base", "red"):
        assert bits(sg["x_" + a], so["x_" + a]) and bits(sg["v_" + a], so["v_" + a]), a
    assert bits(xg, xo) and vclose(vg, vo)


def wide_params(**kw):
    from src_incompressible_openmp_amd.params import poiseuille_params, to_struct

    return to_struct(poiseuille_params(**kw))


def wide_post_collision(oracle, p, seed):
    from src_incompressible_openmp_amd.fileio import uniform_state

    x, v, _, _ = uniform_state(p.n, p.Lx, p.Ly, p.kBT, seed=seed)
    # a checkerboard of occupancy: particles of the "even" cells leave with
    # probability 0.7 for anywhere in the domain, so even cells run low next
    # to fuller odd ones (sparse transfers as well as dense spreads)
    rng = np.random.default_rng(seed)
    px, py = x[0::2], x[1::2]
    even = ((np.floor(px / (p.Lx / p.Nx)) + np.floor(py / (p.Ly / p.Ny))) % 2) == 0
    go = even & (rng.random(p.n) < 0.7)
    px[go] = rng.random(int(go.sum())) * p.Lx
    py[go] = rng.random(int(go.sum())) * p.Ly
    x[0::2], x[1::2] = px, py
    oracle.force_stream(p, x, v)
    oracle.apply_bc(CTR, p, x, v, 104729)
    cid, _ = oracle.shift_cell_ids(CTR, p, x, 130363)
    oracle.collide(CTR, p, v, cid, 433494437)
    return x, v


@pytest.mark.parametrize("passes", [0, 5])
def test_wide_single_zone(gpu, oracle, passes):
    """A zone of 3072 cells: its working set does not fit the 16-bit LDS
    form, so the pass kernel runs with 32-bit indices in global scratch and
    the correction kernel takes its two-sweep form.  Without dense passes
    the sparse phase does all the work."""
    p = wide_params(Nx=64, Ny=48, gamma=8.0, zoneSingleFullDomainMode=True, maxRedistribPasses=passes)
    x, v = wide_post_collision(oracle, p, 11)
    _, rects = gpu.build_zone_layout(p, LAYOUT_BASE)
    assert len(rects) == 1
    xo, vo = x.copy(), v.copy()
    st, _, mv = oracle.zone_kernel(CTR, p, xo, vo, LAYOUT_BASE, 0, 9011)
    xg, vg = x.copy(), v.copy()
    sg, mg = gpu.run_zone_kernel(p, xg, vg, rects[0], 9011)
    print("wide single zone moved", mv, flush=True)
    assert np.array_equal(mg, mv)
    assert bits(xg, xo) and bits(vg, vo)
    assert bits(sg, st)


@pytest.mark.parametrize("layout", [LAYOUT_BASE, LAYOUT_SHIFTED])
def test_wide_zone_pass(gpu, oracle, layout):
    """Zones of 60x48 cells in both layouts (the shifted one overlapping):
    the 32-bit working-set paths of both zone kernels in a snapshot pass."""
    p = wide_params(Nx=120, Ny=48, gamma=6.0, zoneTileNx=60, zoneTileNy=48, maxRedistribPasses=1)
    x, v = wide_post_collision(oracle, p, 12)
    xo, vo = x.copy(), v.copy()
    mo, owo = oracle.zone_pass(CTR, p, xo, vo, layout, 88172645463325252)
    xg, vg = x.copy(), v.copy()
    mg, owg = gpu.zone_pass(p, xg, vg, layout, 88172645463325252)
    print("wide zone pass moved", mo["nParticlesMovedDense"], mo["nParticlesMovedSparse"], flush=True)
    assert np.array_equal(owg, owo)
    assert mg["nParticlesMovedDense"] == mo["nParticlesMovedDense"]
    assert mg["nParticlesMovedSparse"] == mo["nParticlesMovedSparse"]
    assert bits(xg, xo) and bits(vg, vo)
    assert mg["corrRmsDU"] == mo["corrRmsDU"] and mg["corrMaxDU"] == mo["corrMaxDU"]


def clustered_state():
    """64x48 cells, 20 per cell on average, 15% of the particles crowded into
    three columns (about 84 per cell there)."""
    from src_incompressible_openmp_amd.fileio import uniform_state

    p = wide_params(Nx=64, Ny=48, gamma=20.0)
    x, v, _, _ = uniform_state(p.n, p.Lx, p.Ly, p.kBT, seed=21)
    rng = np.random.default_rng(21)
    go = rng.random(p.n) < 0.15
    x[0::2][go] = (20.0 + 3.0 * rng.random(int(go.sum()))) * (p.Lx / p.Nx)
    return p, x, v


def test_full_step_clustered(gpu, oracle):
    """A full step from a state with a crowded band of columns: the binning
    after each zone pass is updated from the one before it (stayers merged
    with arrivals).  The case has runs longer than a register batch, cells
    with many arrivals, and blocks of cells whose runs exceed the LDS staging,
    so every path of the update runs; the step must still equal the oracle's."""
    p, x, v = clustered_state()
    dx = p.Lx / p.Nx
    xo, vo = x.copy(), v.copy()
    so = oracle.step(CTR, p, xo, vo, 3, stages=True)

    def cells(xs):
        ix = np.clip(np.floor(xs[0::2] / dx).astype(np.int64), 0, p.Nx - 1)
        iy = np.clip(np.floor(xs[1::2] / (p.Ly / p.Ny)).astype(np.int64), 0, p.Ny - 1)
        return iy + p.Ny * ix

    c_ref, c_base = cells(so["x_ref"]), cells(so["x_base"])
    moved = c_ref != c_base
    arrivals = np.bincount(c_base[moved], minlength=p.Nx * p.Ny)
    runs = np.bincount(c_ref, minlength=p.Nx * p.Ny)
    block = np.add.reduceat(np.bincount(c_base, minlength=p.Nx * p.Ny), np.arange(0, p.Nx * p.Ny, 128))
    print("clustered: moved", int(moved.sum()), "max arrivals", arrivals.max(), "max run", runs.max(),
          "max block", block.max(), flush=True)
    assert arrivals.max() > 8 and runs.max() > 32 and block.max() > 6144
    sim = gpu.Sim(p)
    sim.set_state(x, v)
    sg = sim.step_stages(3)
    xg, vg = sim.get_state()
    for a in ("ref", "base", "red"):
        assert bits(sg["x_" + a], so["x_" + a]), a
        assert bits(sg["v_" + a], so["v_" + a]), a
    assert bits(xg, xo)
    assert vclose(vg, vo)
    m = sg["metrics"]
    for L in ("base", "shifted"):
        assert m[L]["nParticlesMovedDense"] == so[L]["nParticlesMovedDense"]
        assert m[L]["nParticlesMovedSparse"] == so[L]["nParticlesMovedSparse"]


@pytest.mark.parametrize("case", ["small", "cfg0"])
def test_steps_in_one_call_equal_single_calls(gpu, case):
    """Steps run in one call (each step's kick also streams the step after
    it) equal the same steps run one call each, bit for bit; so do a call
    that reports the last step's metrics and the asynchronous form."""
    p = to_struct(case_params(case))
    x, v, _, _ = case_state(case)
    a = gpu.Sim(p)
    a.set_state(x, v)
    a.step(1, 6)
    xa, va = a.get_state()
    b = gpu.Sim(p)
    b.set_state(x, v)
    for s in range(1, 7):
        b.step(s)
    xb, vb = b.get_state()
    assert bits(xa, xb) and bits(va, vb)
    c = gpu.Sim(p)
    c.set_state(x, v)
    mc = c.step(1, 6, metrics=True)
    d = gpu.Sim(p)
    d.set_state(x, v)
    d.step(1, 5)
    md = d.step(6, metrics=True)
    xc, vc = c.get_state()
    xd, vd = d.get_state()
    assert bits(xc, xa) and bits(vc, va) and bits(xd, xa) and bits(vd, va)
    assert mc == md
    e = gpu.Sim(p)
    e.set_state(x, v)
    e.step_async(1, 6)
    e.check()
    xe, ve = e.get_state()
    assert bits(xe, xa) and bits(ve, va)


def test_clustered_steps_in_one_call(gpu):
    """The clustered state over four steps in one call (each binning after
    streaming updated from the previous step's red binning) equals four
    single-step calls (each binning rebuilt)."""
    p, x, v = clustered_state()
    a = gpu.Sim(p)
    a.set_state(x, v)
    a.step(1, 4)
    b = gpu.Sim(p)
    b.set_state(x, v)
    for s in range(1, 5):
        b.step(s)
    xa, va = a.get_state()
    xb, vb = b.get_state()
    assert bits(xa, xb) and bits(va, vb)


@pytest.mark.parametrize("layout", [LAYOUT_BASE, LAYOUT_SHIFTED])
def test_zone_pass_big_zones(gpu, oracle, layout):
    """25x25-cell zones of about 69,000 particles (110 per cell): the cell
    arrays fit the 16-bit form but the particle count does not, so these
    zones take the 32-bit kernels on their small strided grid while the
    smaller edge zones of the shifted layout take the fused kernel."""
    p = wide_params(Nx=50, Ny=25, gamma=110.0, zoneTileNx=25, zoneTileNy=25, maxRedistribPasses=1)
    x, v = wide_post_collision(oracle, p, 13)
    xo, vo = x.copy(), v.copy()
    mo, owo = oracle.zone_pass(CTR, p, xo, vo, layout, 88172645463325252)
    xg, vg = x.copy(), v.copy()
    mg, owg = gpu.zone_pass(p, xg, vg, layout, 88172645463325252)
    print("big zones moved", mo["nParticlesMovedDense"], mo["nParticlesMovedSparse"], flush=True)
    assert np.array_equal(owg, owo)
    assert mg["nParticlesMovedDense"] == mo["nParticlesMovedDense"]
    assert mg["nParticlesMovedSparse"] == mo["nParticlesMovedSparse"]
    assert bits(xg, xo) and bits(vg, vo)


def test_medium_grid_steps_in_one_call(gpu):
    """256x128 cells, 20 per cell: five steps in one call (fused kick and
    stream, histograms from the kick, binning updates from the passes'
    recorded moves) equal five single-step calls bit for bit."""
    from src_incompressible_openmp_amd.fileio import uniform_state

    p = to_struct(poiseuille_params(Nx=256, Ny=128, gamma=20.0))
    x, v, _, _ = uniform_state(p.n, p.Lx, p.Ly, seed=12)
    a = gpu.Sim(p)
    a.set_state(x, v)
    a.step(1, 5)
    b = gpu.Sim(p)
    b.set_state(x, v)
    for s in range(1, 6):
        b.step(s)
    xa, va = a.get_state()
    xb, vb = b.get_state()
    assert bits(xa, xb) and bits(va, vb)


def test_async_error_reported_once(gpu):
    """A wall-mode error raised inside an asynchronous batch is reported by
    the next check and cleared by it: a later check, and a later clean
    batch, report nothing."""
    prm = case_params("small")
    prm["boundary_bottom"] = "sticky"  # particles reach the bottom wall within a step
    p = to_struct(prm)
    x, v, _, _ = case_state("small")
    sim = gpu.Sim(p)
    try:
        sim.set_state(x, v)
        sim.step_async(1, 1)
        with pytest.raises(gpu.MpcdError) as e:
            sim.check()
        assert e.value.code == -3
        sim.check()  # the flag was cleared by the first check
    finally:
        sim.close()
    prm["boundary_bottom"] = "thermalize"
    sim = gpu.Sim(to_struct(prm))
    try:
        sim.set_state(x, v)
        sim.step_async(1, 2)
        sim.check()
    finally:
        sim.close()


def test_set_state_keeps_the_array_not_uploaded(gpu):
    """After steps have put the slots in cell order, uploading only v (or
    only x) keeps every particle's own position (or velocity)."""
    p = to_struct(case_params("small"))
    x, v, _, _ = case_state("small")
    sim = gpu.Sim(p)
    try:
        sim.set_state(x, v)
        sim.step(1, 2)
        x1, v1 = sim.get_state()
        vnew = v1[::-1].copy()
        sim.set_state(None, vnew)
        x2, v2 = sim.get_state()
        assert bits(x2, x1) and bits(v2, vnew)
        sim.step(3, 1)
        x3, v3 = sim.get_state()
        xnew = x3.copy()
        xnew[0::2] = x3[0::2][::-1]  # particles swap positions, velocities stay theirs
        sim.set_state(xnew, None)
        x4, v4 = sim.get_state()
        assert bits(x4, xnew) and bits(v4, v3)
    finally:
        sim.close()
