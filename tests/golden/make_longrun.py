"""Writes tests/golden/longrun_cfg0.npz: statistics of the compiled
original's benchmark driver (oracle/_ref/main_benchmark_poiseuille_openmp_
profiled_shifted_profiled, built from the reference's own sources by
oracle/Makefile) over R = 8 runs of 1500 steps of the 50x50-cell,
10-particles-per-cell channel from the initial states of seeds 1..8 (see
tests/longrun.py).  The GPU test (tests/test_gpu_longrun.py) runs this
library's driver from the same inputs and compares the statistics.

    python tests/golden/make_longrun.py
"""
from __future__ import annotations

import os
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from tests.longrun import N_RUNS, N_STEPS, cfg0_params, dump_steps, run_set  # noqa: E402

REF_EXE = os.path.join(ROOT, "oracle", "_ref", "main_benchmark_poiseuille_openmp_profiled_shifted_profiled")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "longrun_cfg0.npz")


def main() -> None:
    threads = len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    with tempfile.TemporaryDirectory(prefix="mpcd_longrun_") as d:
        st = run_set(REF_EXE, d, cfg0_params(), threads=threads)
    wall = time.perf_counter() - t0
    np.savez_compressed(OUT, runs=N_RUNS, n_steps=N_STEPS, dumps=np.array(dump_steps()), threads=threads,
                        wall_s=wall, **st)
    print(f"wrote {OUT}: {N_RUNS} runs x {N_STEPS} steps in {wall:.0f} s on {threads} threads")
    for k in ("Qx", "meanKinetic", "occStd", "outBand", "meanVelocityX"):
        print(f"  {k}: {st[k].mean():.6g} +- {st[k].std(ddof=1) / np.sqrt(N_RUNS):.2g}")


if __name__ == "__main__":
    main()
