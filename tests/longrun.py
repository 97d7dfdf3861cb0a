"""Long runs of the benchmark driver and their statistics.

The compiled original's driver
(main_benchmark_poiseuille_openmp_profiled_shifted_profiled, oracle/_ref)
and this library's GPU driver (main_benchmark_poiseuille_gpu) take the same
inputs and write the same outputs, so one function writes the inputs, one
runs either executable and one reads either output into statistics:

* the metrics time series (``_metrics.csv``, columns as
  main_benchmark_poiseuille_openmp_profiled_shifted_profiled.cpp:215; Qx,
  mean kinetic energy, occStd ... as :105-155 compute them);
* the flow profile Ux(y) and the temperature profile T(y): row means over x
  of the ``_cellfields.bin`` dumps (dump_io.cpp:19-39), cell iy + Ny*ix.

The two codes draw different random numbers (the original uses
std::mt19937_64, the GPU the counter-based generator), so they agree only
in distribution.  Each side runs R channels from R initial states; run r of
one side is paired with run r of the other (same initial state), and the
steady-window (second half) means of the pairs' differences are tested
against zero with their spread over the R pairs (a t statistic).  The
slowest flow mode of the channel relaxes over thousands of steps, so a
single run's time average carries no usable error bar; independent runs do.
"""
from __future__ import annotations

import csv
import os
import subprocess
from typing import Dict, List, Sequence

import numpy as np

from src_incompressible_openmp_amd.fileio import read_cell_fields, uniform_state, write_u8, write_xy
from src_incompressible_openmp_amd.params import poiseuille_params, write_params_kv

SERIES = ("Qx", "meanKinetic", "occStd", "outBand", "meanVelocityX")
N_STEPS = 1500
N_RUNS = 8
DUMP_EVERY = 50


def cfg0_params() -> Dict[str, object]:
    """The 50x50-cell, 10-particles-per-cell Poiseuille channel (25,000
    particles, dt 0.05, +-90 deg, body force 1e-3, betaEOS 0.05)."""
    return poiseuille_params(Nx=50, Ny=50, gamma=10.0)


def dump_steps(n_steps: int = N_STEPS, every: int = DUMP_EVERY) -> List[int]:
    """Dumps every `every` steps over the second half of the run."""
    return list(range(n_steps // 2 + every, n_steps + 1, every))


def write_inputs(d: str, prm: Dict[str, object], n_steps: int, dumps: Sequence[int], seed: int) -> None:
    n = int(prm["n"])
    x, v, t, r0 = uniform_state(n, float(prm["Lx"]), float(prm["Ly"]), seed=seed)
    write_params_kv(os.path.join(d, "params.kv"), prm,
                    {"benchmark_nSteps": n_steps, "benchmark_metricsEvery": 1,
                     "benchmark_dumpSteps": ",".join(str(s) for s in dumps) if dumps else "-1"})
    write_xy(os.path.join(d, "x.bin"), x, n)
    write_xy(os.path.join(d, "v.bin"), v, n)
    write_u8(os.path.join(d, "t.bin"), t, n)
    write_xy(os.path.join(d, "r0.bin"), r0, n)


def run_driver(exe: str, d: str, prefix: str = "o", threads: int = 1, timeout: float = 600) -> None:
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    subprocess.run([exe, os.path.join(d, "params.kv"), os.path.join(d, "x.bin"), os.path.join(d, "v.bin"),
                    os.path.join(d, "t.bin"), os.path.join(d, "r0.bin"), os.path.join(d, prefix), "1", "1",
                    str(threads)], check=True, env=env, stdout=subprocess.DEVNULL, timeout=timeout)


def read_metrics(path: str) -> Dict[str, np.ndarray]:
    with open(path) as fh:
        rows = list(csv.DictReader(fh))
    return {k: np.array([float(r[k]) for r in rows]) for k in rows[0]}


def summarize(d: str, prefix: str, n_steps: int, dumps: Sequence[int]) -> Dict[str, np.ndarray]:
    """One run: second-half means of SERIES, every 10th step of Qx, KE and
    occStd, the dump-averaged Ux(y), T(y) and the particle count of every
    dump."""
    m = read_metrics(os.path.join(d, prefix + "_metrics.csv"))
    if not np.array_equal(m["step"], np.arange(1, n_steps + 1)):
        raise AssertionError("metrics rows are not steps 1..n")
    half = n_steps // 2
    out = {k: np.float64(m[k][half:].mean()) for k in SERIES}
    for k in ("Qx", "meanKinetic", "occStd"):
        out[k + "_every10"] = m[k][9::10].copy()
    ux, tt, nn = [], [], []
    for s in dumps:
        f = read_cell_fields(os.path.join(d, f"{prefix}_step{s:04d}_cellfields.bin"))
        Nx, Ny = f["Nx"], f["Ny"]
        ux.append(f["Ux"].reshape(Nx, Ny).mean(axis=0))
        tt.append(f["T"].reshape(Nx, Ny).mean(axis=0))
        nn.append(int(f["N"].sum()))
    out["ux_profile"] = np.mean(ux, axis=0)
    out["t_profile"] = np.mean(tt, axis=0)
    out["n_total"] = np.array(nn, np.int64)
    return out


def run_set(exe: str, workdir: str, prm: Dict[str, object], runs: int = N_RUNS, n_steps: int = N_STEPS,
            threads: int = 1) -> Dict[str, np.ndarray]:
    """R runs from initial-state seeds 1..R; arrays stacked over runs."""
    dumps = dump_steps(n_steps)
    per: List[Dict[str, np.ndarray]] = []
    for r in range(runs):
        d = os.path.join(workdir, f"run{r}")
        os.makedirs(d, exist_ok=True)
        write_inputs(d, prm, n_steps, dumps, seed=r + 1)
        run_driver(exe, d, threads=threads)
        per.append(summarize(d, "o", n_steps, dumps))
        for f in os.listdir(d):
            if f.startswith("o_step"):
                os.remove(os.path.join(d, f))
    return {k: np.array([p[k] for p in per]) for k in per[0]}


def compare(gpu: Dict[str, np.ndarray], ref: Dict[str, np.ndarray], t_max: float = 5.0,
            profile_rms: float = 2.0) -> Dict[str, float]:
    """t statistics of the paired differences (gpu run r - ref run r) of the
    window means; for each profile, per row, with their max and rms over the
    rows.  Raises AssertionError past t_max (rms past profile_rms)."""
    R = len(ref["Qx"])
    assert len(gpu["Qx"]) == R
    scores: Dict[str, float] = {}

    def tstat(d: np.ndarray) -> np.ndarray:
        se = d.std(axis=0, ddof=1) / np.sqrt(R)
        return np.abs(d.mean(axis=0)) / np.where(se > 0, se, np.inf)

    for k in SERIES:
        scores[k] = float(tstat(gpu[k] - ref[k]))
        scores[k + "_rel"] = float(abs(np.mean(gpu[k] - ref[k])) / max(abs(float(np.mean(ref[k]))), 1e-300))
    for k in ("ux_profile", "t_profile"):
        t = tstat(gpu[k] - ref[k])
        scores[k + "_tmax"] = float(t.max())
        scores[k + "_trms"] = float(np.sqrt(np.mean(t ** 2)))
    bad = {k: v for k, v in scores.items()
           if (k.endswith("_trms") and v > profile_rms) or (k.endswith("_tmax") and v > t_max + 2.0) or
           (k in SERIES and v > t_max)}
    assert not bad, f"long-run statistics differ: {bad} (all: {scores})"
    return scores
