"""Batched ensembles on the GPU: every replica of an ensemble of uneven
channels (grid sizes, particle densities, forcings, walls, zone tiles, seed
keys) against a single simulation of its own parameters with the same seed
key, bit for bit, over steps run eagerly and as replayed HIP graphs of both
buffer parities; the per-replica flow-rate sums against the single
simulation's Qx; and the refused configurations.

The single simulation is itself pinned to the oracle (test_gpu_parity.py,
test_gpu_scale.py); a seed key only offsets the seeds (key 0 is the
benchmark driver's own), which the replica with key 0 checks directly."""
import numpy as np
import pytest

from src_incompressible_openmp_amd.fileio import uniform_state
from src_incompressible_openmp_amd.params import poiseuille_params, to_struct
from tests.golden.cases import case_params
from tests.oracle_lib import CTR

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    from src_incompressible_openmp_amd import gpu as g

    g.load()
    assert g.device_count() >= 1, "no GPU visible"
    return g


def replicas():
    bcmix = dict(case_params("bcmix"), keepMeanFlow=False)
    prms = [case_params("small"),
            poiseuille_params(Nx=24, Ny=12, gamma=6.0, bodyForceX=0.003),
            case_params("override"),
            bcmix,
            poiseuille_params(Nx=50, Ny=50, gamma=10.0),
            poiseuille_params(Nx=37, Ny=21, gamma=13.0, bodyForceX=2e-4, zoneTileNx=9, zoneTileNy=7),
            poiseuille_params(Nx=160, Ny=90, gamma=7.0, bodyForceX=5e-3)]  # 28 base zones of the default tile
    keys = [0, 7, 123, 5, 99, 2 ** 40 + 3, 31]
    return [to_struct(p) for p in prms], keys


def states(ps):
    return [uniform_state(p.n, p.Lx, p.Ly, seed=11 + r)[:2] for r, p in enumerate(ps)]


def single_run(gpu, p, key, x, v, steps):
    sim = gpu.Sim(p)
    try:
        sim.set_seed_key(key)
        sim.set_state(x, v)
        sim.step(1, steps)
        return sim.get_state()
    finally:
        sim.close()


def test_replicas_match_single_simulations(gpu):
    ps, keys = replicas()
    st = states(ps)
    ens = gpu.Ensemble(ps, keys)
    try:
        info = ens.info()
        assert info["particles"] == sum(p.n for p in ps) and info["cells"] == sum(p.Nc for p in ps)
        for r, (x, v) in enumerate(st):
            ens.set_state(r, x, v)
        ens.step(1, 3, graph=False)  # eager
        ens.step(4, 3, graph=True)   # graphs of both parities, captured and replayed
        ens.step(7, 2, graph=True)
        got = [ens.get_state(r) for r in range(len(ps))]
    finally:
        ens.close()
    for r, p in enumerate(ps):
        print(f"ensemble replica {r}: {p.Nx}x{p.Ny}, n={p.n}, key {keys[r]}", flush=True)
        xs, vs = single_run(gpu, p, keys[r], st[r][0], st[r][1], 8)
        assert got[r][0].tobytes() == xs.tobytes(), r
        assert got[r][1].tobytes() == vs.tobytes(), r


def test_graph_replays_match_eager_steps(gpu):
    """The same ensemble stepped eagerly and by graph replays, with host
    work and synchronisations between the replays: every replay's state
    equals the eager one."""
    ps, keys = replicas()
    st = states(ps)
    runs = []
    for use_graph in (False, True):
        ens = gpu.Ensemble(ps, keys)
        try:
            for r, (x, v) in enumerate(st):
                ens.set_state(r, x, v)
            snaps = []
            for s in range(1, 6):
                ens.step(s, 1, graph=use_graph)
                snaps.append([ens.get_state(r) for r in (0, len(ps) - 1)])
            for s in (6, 8):  # two steps per call: the first kick streams the second step
                ens.step(s, 2, graph=use_graph)
                snaps.append([ens.get_state(r) for r in (0, len(ps) - 1)])
            runs.append(snaps)
        finally:
            ens.close()
    for s, (a, b) in enumerate(zip(*runs)):
        for (xa, va), (xb, vb) in zip(a, b):
            assert xa.tobytes() == xb.tobytes() and va.tobytes() == vb.tobytes(), s


def test_seed_key_zero_is_the_oracle_step(gpu, oracle):
    """Key 0 draws with the benchmark driver's seeds: replica 0 after one
    step equals the oracle's step."""
    ps, keys = replicas()
    st = states(ps)
    ens = gpu.Ensemble(ps[:2], keys[:2])
    try:
        for r in range(2):
            ens.set_state(r, *st[r])
        ens.step(1, 1, graph=False)
        xg, vg = ens.get_state(0)
    finally:
        ens.close()
    xo, vo = st[0][0].copy(), st[0][1].copy()
    oracle.step(CTR, ps[0], xo, vo, 1)
    assert xg.tobytes() == xo.tobytes()
    assert float(np.max(np.abs(vg - vo))) <= 1e-12 * max(1.0, float(np.max(np.abs(vo))))


def test_flow_rates_match_single_simulations(gpu):
    ps, keys = replicas()
    st = states(ps)
    ens = gpu.Ensemble(ps, keys)
    try:
        for r, (x, v) in enumerate(st):
            ens.set_state(r, x, v)
        ens.flow(True)
        ens.step(1, 4, graph=True)
        q, q2, ns = ens.flow(False, fetch=True)
    finally:
        ens.close()
    assert ns == 4
    for r, p in enumerate(ps):
        sim = gpu.Sim(p)
        try:
            sim.set_seed_key(keys[r])
            sim.set_state(*st[r])
            qs = [sim.step_diag(s)[1]["Qx"] for s in range(1, 5)]
        finally:
            sim.close()
        tol = 1e-10 * max(1e-12, sum(abs(a) for a in qs))
        assert abs(q[r] - sum(qs)) <= tol, (r, q[r], sum(qs))
        assert abs(q2[r] - sum(a * a for a in qs)) <= 1e-10 * max(1e-24, sum(a * a for a in qs)), r


def test_refused_configurations(gpu):
    ps, keys = replicas()
    bad = to_struct(dict(case_params("bcmix")))  # keepMeanFlow
    with pytest.raises(gpu.MpcdError):
        gpu.Ensemble([ps[0], bad], [0, 1])
    with pytest.raises(gpu.MpcdError):
        gpu.Ensemble([to_struct(dict(case_params("small"), useLiquidClosure=False))])
