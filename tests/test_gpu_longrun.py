"""Long-run statistical parity with the compiled original.

Eight runs of 1500 steps of the 50x50-cell, 10-particles-per-cell channel
(25,000 particles) through this library's GPU driver, from the same eight
initial states as the compiled original's runs recorded in
tests/golden/longrun_cfg0.npz (tests/golden/make_longrun.py).  The two
codes draw different random numbers, so they are compared in distribution
(tests/longrun.py): the paired differences of the second-half means of Qx,
the mean kinetic energy, occStd, outBand and the mean velocity, and of the
dump-averaged Ux(y) and T(y) profiles, must be zero within their spread
over the runs (|t| <= 5; profiles: rms t over the rows <= 2).  The particle
count of every dump is exact."""
import os

import numpy as np
import pytest

from tests.longrun import N_RUNS, N_STEPS, SERIES, cfg0_params, compare, run_set

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "longrun_cfg0.npz")


@pytest.mark.timeout(600)
def test_longrun_statistics_match_original(tmp_path):
    from src_incompressible_openmp_amd import build, gpu

    gpu.load()
    ref = dict(np.load(GOLDEN))
    assert int(ref["runs"]) == N_RUNS and int(ref["n_steps"]) == N_STEPS
    prm = cfg0_params()
    got = run_set(build.DRIVER, str(tmp_path), prm)
    assert np.all(got["n_total"] == int(prm["n"]))
    assert np.all(ref["n_total"] == int(prm["n"]))
    scores = compare(got, ref)
    print("longrun scores:", {k: round(v, 3) for k, v in scores.items()}, flush=True)
    for k in SERIES:
        print(f"  {k}: gpu {got[k].mean():.6g} ref {ref[k].mean():.6g}", flush=True)
