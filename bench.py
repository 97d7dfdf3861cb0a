"""Benchmark: particle-steps/s of the SRD/MPCD time step on one MI355X.

Workload: 2-D periodic-x Poiseuille channel, 1024 x 512 cells of size 0.2
(Lx = 204.8, Ly = 102.4), 20 particles per cell (10,485,760 particles), FP64
structure-of-arrays; dt 0.05, +-90 deg rotation, body force 1e-3, thermalising
walls, base + shifted zone passes and the liquid closure every step.  The
initial state is synthetic (uniform positions, Maxwell velocities) and is
resident in HBM before the timed region.

Prints one JSON line (contract in the module's CLI help).  ``roofline``
describes the dominant kernel: algorithmic bytes per launch over its average
launch time, measured with HIP events on the simulation's own stream during
the timed steps.  ``cpu_baseline`` times the compiled original reference
(oracle/_ref, OpenMP) on a bounded sample of the same physics.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from src_incompressible_openmp_amd import gpu  # noqa: E402
from src_incompressible_openmp_amd.fileio import uniform_state, write_u8, write_xy  # noqa: E402
from src_incompressible_openmp_amd.params import poiseuille_params, to_struct, write_params_kv  # noqa: E402

METRIC = "particle-steps/s plain vs learned-closure at fixed Poiseuille profile error"
HBM_PEAK_GBS = 8000.0
# algorithmic FP64 bytes per particle of each kernel class (launch covers all particles)
ALG_BYTES = {"stream_bc": 64, "collide": 48, "zone_kernel": 48, "kick": 48}
STEP_ALG_BYTES = 256  # compulsory particle traffic of one plain step


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(nx: int, ny: int, ppc: int, steps: int):
    """The compiled original's benchmark driver on a bounded sample."""
    exe = os.path.join(ROOT, "oracle", "_ref", "main_benchmark_poiseuille_openmp_profiled_shifted_profiled")
    if not os.path.exists(exe):
        return None
    threads = max(1, min(16, os.cpu_count() or 1))
    prm = poiseuille_params(Nx=nx, Ny=ny, gamma=float(ppc))
    n = int(prm["n"])
    x, v, t, r0 = uniform_state(n, float(prm["Lx"]), float(prm["Ly"]), seed=3)
    d = tempfile.mkdtemp(prefix="mpcd_cpu_")
    try:
        write_params_kv(os.path.join(d, "params.kv"), prm,
                        {"benchmark_nSteps": steps, "benchmark_metricsEvery": 0, "benchmark_dumpSteps": "-1"})
        write_xy(os.path.join(d, "x.bin"), x, n)
        write_xy(os.path.join(d, "v.bin"), v, n)
        write_u8(os.path.join(d, "t.bin"), t, n)
        write_xy(os.path.join(d, "r0.bin"), r0, n)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        t0 = time.perf_counter()
        subprocess.run([exe, os.path.join(d, "params.kv"), os.path.join(d, "x.bin"), os.path.join(d, "v.bin"),
                        os.path.join(d, "t.bin"), os.path.join(d, "r0.bin"), os.path.join(d, "o"), "1", "1",
                        str(threads)], check=True, env=env, stdout=subprocess.DEVNULL, timeout=600)
        wall = time.perf_counter() - t0
        kv = {}
        with open(os.path.join(d, "o_runout.kv")) as fh:
            for line in fh:
                if "=" in line:
                    k, val = line.strip().split("=", 1)
                    kv[k] = val
        per_step = float(kv["timePerStepMean"])
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return {
        "value": n / per_step, "unit": "particle-steps/s", "cores": threads, "kind": "reference",
        "sample": (f"original OpenMP driver (main_benchmark_poiseuille_openmp_profiled_shifted_profiled, g++ -O3 "
                   f"-fopenmp) on {nx}x{ny} cells x {ppc} ppc (n={n}), {steps} steps, same physics; "
                   f"its cost per step grows ~quadratically with domain size, so at 1024x512 its throughput is "
                   f"lower than this sample's"),
        "s_per_step": per_step, "wall_s": wall,
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--ppc", type=int, default=20)
    ap.add_argument("--cpu-nx", type=int, default=256)
    ap.add_argument("--cpu-ny", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true", help="time without per-kernel events")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU decomposition is described in DESIGN.md, not implemented)", file=sys.stderr)
        return 2

    prm = poiseuille_params(Nx=args.nx, Ny=args.ny, gamma=float(args.ppc))
    p = to_struct(prm)
    n = p.n
    log(f"bench: {args.nx}x{args.ny} cells, {args.ppc} ppc, n={n}; building state")
    x, v, _, _ = uniform_state(n, p.Lx, p.Ly, seed=1)
    sim = gpu.Sim(p)
    sim.set_state(x, v)
    del x, v
    log(f"bench: warmup {args.warmup} steps")
    sim.step(1, args.warmup)
    gpu.device_synchronize()
    if not args.no_profile:
        sim.profile(True)
    log(f"bench: timing {args.steps} steps")
    t0 = time.perf_counter()
    sim.step_async(1 + args.warmup, args.steps)
    sim.synchronize()
    gpu.device_synchronize()
    t1 = time.perf_counter()
    sim.check()
    elapsed = t1 - t0
    prof = None if args.no_profile else sim.profile_read()
    sim.profile(False)
    m = sim.step(1 + args.warmup + args.steps, 1, metrics=True)
    sim.close()

    value = n * args.steps / elapsed
    ms_per_step = 1e3 * elapsed / args.steps
    out = {
        "metric": METRIC, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"poiseuille_{args.nx}x{args.ny}_ppc{args.ppc}", "closure": "plain",
                   "cells": args.nx * args.ny, "particles": n, "dt": p.dt, "alpha_deg": 90.0,
                   "body_force": p.bodyForceX, "betaEOS": p.betaEOS, "zone_tile": [p.zoneTileNx, p.zoneTileNy]},
        "step_roofline_frac": value * STEP_ALG_BYTES / (HBM_PEAK_GBS * 1e9),
    }
    if prof:
        per = {k: {"ms_total": t, "launches": c, "ms_per_launch": (t / c if c else 0.0)} for k, (t, c) in prof.items()}
        out["kernel_ms_per_step"] = {k: per[k]["ms_total"] / args.steps for k in per}
        cands = {k: per[k]["ms_total"] for k in ALG_BYTES if per[k]["launches"]}
        dom = max(cands, key=cands.get)
        dur = per[dom]["ms_per_launch"] * 1e-3
        alg = ALG_BYTES[dom] * n
        ach = alg / dur / 1e9
        out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": ach / HBM_PEAK_GBS, "traffic": None,
                           "alg_bytes_per_launch": alg, "ms_per_launch": per[dom]["ms_per_launch"]}
    out["last_step"] = {"moved_dense": m["base"]["nParticlesMovedDense"] + m["shifted"]["nParticlesMovedDense"],
                        "moved_sparse": m["base"]["nParticlesMovedSparse"] + m["shifted"]["nParticlesMovedSparse"],
                        "beta_applied": m["closure"]["betaRepairApplied"]}
    if not args.no_cpu_baseline:
        log(f"bench: CPU baseline {args.cpu_nx}x{args.cpu_ny} x {args.ppc}, {args.cpu_steps} steps")
        cb = cpu_baseline(args.cpu_nx, args.cpu_ny, args.ppc, args.cpu_steps)
        out["cpu_baseline"] = cb
        if cb:
            out["gpu_over_cpu_sample"] = value / cb["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
