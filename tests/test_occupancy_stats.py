"""The oracle's occupancy statistics of a zone pass (occupancy_stats,
restating zone_pass.cpp:975-1044) against the values the compiled original
wrote in its Gauss-Seidel pass runout (tests/golden, 6 significant digits).
The GPU reduction is compared with this oracle in tests/test_gpu_diag.py."""
import numpy as np
import pytest

from src_incompressible_openmp_amd.params import to_struct
from tests.golden.cases import SMALL_CASES, case_params
from tests.test_oracle_golden import close6, kv, load


def counts(p, x):
    ix = np.clip(np.floor(x[0::2] / (p.Lx / p.Nx)).astype(np.int64), 0, p.Nx - 1)
    iy = np.clip(np.floor(x[1::2] / (p.Ly / p.Ny)).astype(np.int64), 0, p.Ny - 1)
    return np.bincount(iy + p.Ny * ix, minlength=p.Nc).astype(np.int32)


@pytest.mark.parametrize("case", SMALL_CASES)
@pytest.mark.parametrize("layout", ["base", "shifted"])
def test_pass_occupancy_stats_match_original(oracle, case, layout):
    g = load(case)
    p = to_struct(case_params(case))
    r = kv(g, f"s2{layout}")
    for tag, x in (("Before", g["x0"]), ("After", g[f"s2{layout}_x"])):
        s = oracle.occupancy_stats(p, counts(p, x))
        assert close6(s["occStd"], r["occStd" + tag]), tag
        assert close6(s["outBand"], r["outBand" + tag]), tag
        assert s["nDense"] == r["nDense" + tag] and s["nSparse"] == r["nSparse" + tag], tag
        assert s["nEmptyActive"] == r["nEmptyActive" + tag], tag


def test_occupancy_closed_forms(oracle):
    """Uniform counts: occStd 0 and nothing out of band; one cell raised by
    k over a uniform grid of Nc cells: occStd = k sqrt(Nc-1)/Nc/gamma."""
    p = to_struct(case_params("small"))
    g = int(p.gamma)
    c = np.full(p.Nc, g, np.int32)
    s = oracle.occupancy_stats(p, c)
    assert s["occStd"] == 0.0 and s["outBand"] == 0.0 and s["nDense"] == 0 and s["nSparse"] == 0
    k = 7
    c[3] += k
    s = oracle.occupancy_stats(p, c)
    want = k * np.sqrt(p.Nc - 1) / p.Nc / p.gamma
    assert abs(s["occStd"] - want) <= 1e-14 * want
    assert s["nDense"] == 1 and s["outBand"] == 1.0 / p.Nc


def test_local_fluid_fraction_thresholds_change_empty_cells(oracle):
    """With local thresholds an empty cell next to a full region gets the
    smoothed (not the constant) empty-cell thresholds."""
    prm = case_params("small")
    prm["useLocalFluidFractionThresholds"] = True
    p = to_struct(prm)
    c = np.full(p.Nc, int(p.gamma), np.int32)
    c[0] = 0
    s = oracle.occupancy_stats(p, c)
    assert s["nEmptyActive"] == 1 and s["nDense"] == 0
