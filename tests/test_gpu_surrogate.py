"""Closure surrogate kernels against the CPU reference (tests/surrogate_ref.py,
which rounds to bfloat16 at the same points as the GPU and sums in float64):
the matrix-core lane layout, forward, loss and gradient, the Adam step,
training on the GPU, and the surrogate in the resident step."""
import numpy as np
import pytest
import torch

from tests import surrogate_ref as R

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    from src_incompressible_openmp_amd import gpu, surrogate

    gpu.load()
    assert gpu.device_count() >= 1
    return surrogate


def to_gpu(t):
    """[B,C,X,Y] -> [B, X*Y, C] (cell index iy + Ny*ix)"""
    t = np.asarray(t)
    B, C, X, Y = t.shape
    return np.ascontiguousarray(t.transpose(0, 2, 3, 1).reshape(B, X * Y, C))


def from_gpu(a, X, Y):
    B, _, C = a.shape
    return a.reshape(B, X, Y, C).transpose(0, 3, 1, 2)


def bf16_bits(x):
    t = torch.as_tensor(np.asarray(x, np.float32)).to(torch.bfloat16)
    return t.view(torch.int16).numpy().astype(np.uint16), t.to(torch.float64).numpy()


def test_mfma_lane_layout(S):
    """A[i][k] from lane 16*(k//8) + i, element k%8; B[k][n] from lane
    16*(k//8) + n; D[i][n] to lane 16*(i//4) + n, element i%4."""
    rng = np.random.default_rng(0)
    A = rng.normal(size=(16, 32))
    Bm = rng.normal(size=(32, 16))
    a = np.zeros((64, 8))
    b = np.zeros((64, 8))
    for lane in range(64):
        for j in range(8):
            a[lane, j] = A[lane % 16, 8 * (lane // 16) + j]
            b[lane, j] = Bm[8 * (lane // 16) + j, lane % 16]
    abits, aval = bf16_bits(a)
    bbits, bval = bf16_bits(b)
    Aq = np.zeros((16, 32))
    Bq = np.zeros((32, 16))
    for lane in range(64):
        for j in range(8):
            Aq[lane % 16, 8 * (lane // 16) + j] = aval[lane, j]
            Bq[8 * (lane // 16) + j, lane % 16] = bval[lane, j]
    want = Aq @ Bq
    d = S.mfma_probe(abits, bbits)
    got = np.zeros((16, 16))
    for lane in range(64):
        for r in range(4):
            got[4 * (lane // 16) + r, lane % 16] = d[lane, r]
    assert np.max(np.abs(got - want)) <= 1e-4 * np.max(np.abs(want))


GRIDS = [(12, 10, True, False), (9, 7, False, True), (8, 16, True, True)]


def rand_case(seed, X, Y, B=2):
    rng = np.random.default_rng(seed)
    f = rng.normal(size=(B, 4, X, Y)) * np.array([0.3, 0.5, 0.5, 0.2]).reshape(1, 4, 1, 1)
    t = rng.normal(size=(B, 2, X, Y))
    return f, t


@pytest.mark.parametrize("X,Y,perx,pery", GRIDS)
def test_forward_matches_reference(S, X, Y, perx, pery):
    theta = R.init_params(11)
    f, _ = rand_case(1, X, Y)
    net = S.Net(X, Y, perx, pery, max_batch=2)
    net.set_params(theta)
    got = from_gpu(net.forward(to_gpu(f)), X, Y)
    want, _ = R.forward(theta, f.astype(np.float32), perx, pery, bf=True)
    want = want.numpy()
    assert np.max(np.abs(got - want)) <= 5e-3 * np.max(np.abs(want))


@pytest.mark.parametrize("X,Y,perx,pery", GRIDS)
def test_loss_and_gradient_match_reference(S, X, Y, perx, pery):
    theta = R.init_params(12)
    f, t = rand_case(2, X, Y)
    net = S.Net(X, Y, perx, pery, max_batch=2)
    net.set_params(theta)
    loss, g = net.loss_grad(to_gpu(f), to_gpu(t))
    lref, gref, _ = R.loss_and_grad(theta, f.astype(np.float32), t.astype(np.float32), perx, pery, bf=True)
    assert abs(loss - lref) <= 1e-3 * lref
    assert np.linalg.norm(g - gref) <= 2e-2 * np.linalg.norm(gref)
    # per layer, so that no layer's gradient hides behind another's size
    o = 0
    for co, _, ci in R.layer_shapes():
        k = co * 9 * ci + co
        assert np.linalg.norm(g[o:o + k] - gref[o:o + k]) <= 3e-2 * np.linalg.norm(gref[o:o + k]), (co, ci)
        o += k


def test_adam_step_matches_reference(S):
    theta = R.init_params(13)
    net = S.Net(6, 5, True, False)
    net.set_params(theta)
    rng = np.random.default_rng(3)
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)
    th = theta.copy()
    for step in (1, 2, 3):
        g = rng.normal(size=theta.size)
        net.adam_step(1e-3, g)
        th, m, v = R.adam(th, m, v, g, step, 1e-3)
    assert np.max(np.abs(net.get_params() - th)) <= 1e-12


def test_training_fits_a_teacher(S):
    """Targets from a teacher network; the student (other initial weights)
    trained by the GPU's own kernels lowers the loss many times over."""
    X, Y = 16, 12
    rng = np.random.default_rng(5)
    feats = rng.normal(size=(16, 4, X, Y)) * 0.5
    teacher = R.init_params(21) * 0.5
    tgt = R.forward(teacher, feats.astype(np.float32), True, False, bf=False)[0].numpy()
    net = S.Net(X, Y, True, False, max_batch=4)
    net.set_params(R.init_params(22))
    losses = net.train(to_gpu(feats), to_gpu(tgt), batch=4, iters=300, lr=2e-3, seed=7)
    assert np.all(np.isfinite(losses))
    assert losses[-20:].mean() < 0.2 * losses[:5].mean()


def closure_run_params(Nx=48, Ny=24, gamma=10.0):
    from src_incompressible_openmp_amd.params import poiseuille_params, to_struct

    return to_struct(poiseuille_params(Nx=Nx, Ny=Ny, gamma=gamma, zoneTileNx=12, zoneTileNy=12))


def test_surrogate_learns_the_closure_and_steps(S):
    """Samples from a plain run train the network on the GPU; on held-out
    samples it explains a clear share of the closure's correction (what the
    network gains over no correction); a run that uses it keeps the particle
    count, kinetic energy and flow of the plain run."""
    from src_incompressible_openmp_amd import gpu
    from src_incompressible_openmp_amd.fileio import uniform_state

    p = closure_run_params()
    x, v, _, _ = uniform_state(p.n, p.Lx, p.Ly, seed=3)
    sim = gpu.Sim(p)
    sim.set_state(x, v)
    sim.step(1, 40)
    sim.closure_sample(True, 1.0)
    feats, tgts = [], []
    for s in range(41, 161):
        sim.step(s)
        f, t = sim.closure_sample(True, 1.0, fetch=True)
        feats.append(f)
        tgts.append(t)
    sim.closure_sample(False)
    feats = np.stack(feats)
    tgts = np.stack(tgts)
    scale = float(np.std(tgts))
    tgts = tgts / scale
    train, test = slice(0, 100), slice(100, 120)
    net = S.Net(p.Nx, p.Ny, True, False, max_batch=8)
    net.set_params(R.init_params(1))
    losses = net.train(feats[train], tgts[train], batch=8, iters=400, lr=2e-3, seed=3)
    assert np.all(np.isfinite(losses))
    pred = net.forward(feats[test])
    resid = np.mean((pred - tgts[test]) ** 2)
    base = np.mean(tgts[test] ** 2)
    print(f"surrogate: train loss {losses[:5].mean():.3g} -> {losses[-20:].mean():.3g}, held-out R2 {1 - resid / base:.3f}")
    # the target's repair part follows the zone passes' random moves, which
    # the red state's fields do not show; the EOS part is smooth and
    # learnable: the network explains a clear share of the correction
    assert resid < 0.85 * base
    # the resident loop with the surrogate (plain closure every 10th step)
    xs, vs = sim.get_state()
    ref = gpu.Sim(p)
    ref.set_state(xs, vs)
    ref.step(161, 60)
    xr, vr = ref.get_state()
    lrn = gpu.Sim(p)
    lrn.set_state(xs, vs)
    lrn.set_surrogate(net, 10, scale, 50.0 * scale)
    lrn.step(161, 60)
    xl, vl = lrn.get_state()
    c = lrn.surrogate_counts()
    assert c["learned"] + c["fallback"] == 54 and c["plain"] == 6
    assert xl.shape == xr.shape and np.all(np.isfinite(xl)) and np.all(np.isfinite(vl))
    ke_r, ke_l = 0.5 * np.mean(vr ** 2) * 2, 0.5 * np.mean(vl ** 2) * 2
    assert abs(ke_l - ke_r) <= 0.05 * ke_r
    # mean flow: the two runs differ by thermal noise of the wall exchanges
    assert abs(np.mean(vl[0::2]) - np.mean(vr[0::2])) <= 0.03


def test_failed_residual_checks_give_the_plain_run(S):
    """The residual check decides on the device and a failed step runs the
    plain closure's kernels (predicated on the check's flag) on the same
    fields: a surrogate whose every prediction fails (beyond the cap, or
    tolerances too tight to pass) gives the plain run bit for bit, with every
    learned step counted as a fallback."""
    from src_incompressible_openmp_amd import gpu
    from src_incompressible_openmp_amd.fileio import uniform_state

    p = closure_run_params()
    x, v, _, _ = uniform_state(p.n, p.Lx, p.Ly, seed=5)
    ref = gpu.Sim(p)
    ref.set_state(x, v)
    ref.step(5, 30)  # steps 5..34: the first is a plain one, so the check has its reference
    xr, vr = ref.get_state()
    net = S.Net(p.Nx, p.Ny, True, False, max_batch=1)
    net.set_params(R.init_params(2))
    for cap, occ_tol, div_tol in ((1e-30, 10.0, 10.0), (1e3, 1e-9, 1e-9)):
        lrn = gpu.Sim(p)
        lrn.set_state(x, v)
        lrn.set_surrogate(net, 5, 0.01, cap)
        lrn.set_surrogate_check(occ_tol, div_tol)
        lrn.step(5, 30)
        xl, vl = lrn.get_state()
        c = lrn.surrogate_counts()
        assert c == {"learned": 0, "plain": 6, "fallback": 24}, c
        assert xl.tobytes() == xr.tobytes() and vl.tobytes() == vr.tobytes()
        lrn.close()
    # loose tolerances: a small, finite prediction passes
    lrn = gpu.Sim(p)
    lrn.set_state(x, v)
    lrn.set_surrogate(net, 5, 1e-4, 1e3)
    lrn.set_surrogate_check(1e3, 1e3)
    lrn.step(5, 30)
    c = lrn.surrogate_counts()
    assert c["learned"] == 24 and c["fallback"] == 0, c
