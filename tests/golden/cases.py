"""Golden cases: parameter sets and inputs shared by the fixture generator
(make_goldens.py) and the tests that read the fixtures."""
from __future__ import annotations

import math
from typing import Dict

import numpy as np

from src_incompressible_openmp_amd.fileio import uniform_state
from src_incompressible_openmp_amd.params import poiseuille_params


def case_params(name: str) -> Dict[str, object]:
    if name == "small":
        # 20x10 cells, 10 per cell, 5x5 tiles -> 8 base / 15 shifted zones
        return poiseuille_params(Nx=20, Ny=10, gamma=10.0, zoneTileNx=5, zoneTileNy=5)
    if name == "bcmix":
        # every wall mode, local fluid-fraction thresholds, "1+coef" modes,
        # mean-flow removal; x not periodic
        return poiseuille_params(
            Nx=12, Ny=10, gamma=6.0, zoneTileNx=4, zoneTileNy=4,
            boundary_left="thermalize", boundary_right="bounceback",
            boundary_bottom="specular", boundary_top="bounceback", Utop=0.3, Ubottom=-0.2, wallSigma=1.3,
            useLocalFluidFractionThresholds=True, highMode="1+coef", lowMode="1+coef", coef=0.2,
            lowThrFloor=3.0, highThrFloor=8.0, keepMeanFlow=True, dt=0.08, bodyForceX=0.02, g=-0.01,
            alpha=math.radians(130.0))
    if name == "override":
        # lowThrBulkOverride, thermostat off, periodic y, fixed beta repair
        return poiseuille_params(
            Nx=15, Ny=15, gamma=8.0, zoneTileNx=6, zoneTileNy=7, boundary_bottom="periodic",
            boundary_top="periodic", lowThrBulkOverride=6.5, useThermostat=False, useOptimalBetaRepair=False,
            betaRepair=0.7, maxRedistribPasses=3, smoothPdriveInterzoneWidthCells=1,
            smoothPdriveInterzoneBlend=0.6, smoothPdriveIncludeShiftedLayouts=False)
    if name == "single":
        # one zone covering the domain: snapshot and Gauss-Seidel passes agree
        return poiseuille_params(Nx=16, Ny=8, gamma=8.0, zoneSingleFullDomainMode=True)
    if name == "cfg0":
        # the 50x50, 10-per-cell channel (n = 25,000), 25x25 tiles
        return poiseuille_params(Nx=50, Ny=50, gamma=10.0)
    raise KeyError(name)


SMALL_CASES = ("small", "bcmix", "override", "single")
DIGEST_CASES = ("cfg0",)


def case_state(name: str):
    p = case_params(name)
    seed = {"small": 1, "bcmix": 2, "override": 3, "single": 4, "cfg0": 1}[name]
    x, v, t, r0 = uniform_state(int(p["n"]), float(p["Lx"]), float(p["Ly"]), float(p["kBT"]), seed=seed)
    if name == "bcmix":
        v *= 2.5  # more wall crossings per step
    return x, v, t, r0


def ref_red_pair(x: np.ndarray, v: np.ndarray, Lx: float, Ly: float, seed: int = 7, move_frac: float = 0.08):
    """An ad-hoc (ref, red) pair for the closure: a fraction of the particles
    relocated uniformly with a velocity bias (the idea of
    build_ad_hoc_ref_red_states.m:28-67; own generator)."""
    rng = np.random.default_rng(seed)
    n = x.size // 2
    xr, vr = x.copy(), v.copy()
    idx = rng.choice(n, size=max(1, int(move_frac * n)), replace=False)
    xr[2 * idx] = Lx * rng.random(idx.size)
    xr[2 * idx + 1] = Ly * rng.random(idx.size)
    vr[2 * idx] += 0.3
    vr[2 * idx + 1] -= 0.1
    return xr, vr
