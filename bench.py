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
as many steps again right after the timed ones (an event between kernels
leaves the stream a gap, so ``value`` is timed without them;
``profiled_ms_per_step`` gives the step time with them).  ``cpu_baseline`` times the compiled original reference
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
# compulsory bytes per particle of each kernel class (a launch covers all particles):
# stream_bc x,y,vx,vy read+write; collide vx,vy read+write + slot list (x2: in, out)
# as staged; zone_kernel (the passes and the correction in one kernel, one
# launch per tile pass): the tile pass's compulsory particle traffic, x,y,vx,vy
# read and vx,vy written (48 B; the phase's figure in the step's 256 B);
# kick x,y,vx,vy,pid and the slot list read, x,y,vx,vy,pid written in red-cell
# order (76 B); kick_stream (the kick fused with the next step's stream, all
# steps of a call but its last) the same plus both cell ids written (84 B)
ALG_BYTES = {"stream_bc": 64, "collide": 48, "zone_kernel": 48, "kick": 76, "kick_stream": 84}
STEP_ALG_BYTES = 256  # compulsory particle traffic of one plain step
# kernel of each timed class, as rocprofv3 names it
KERNEL_NAME = {"stream_bc": "k_stream_bc", "collide": "k_collide", "zone_kernel": "k_zone_fused<unsigned short, false>",
               "kick": "k_kick_reorder", "kick_stream": "k_kick_reorder_stream"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pmc_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE, WRITE_SIZE per dispatch, each scaled by its copy
    calibration factor), or None when there is none for this workload."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.md")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 6 and cells[0] == kernel:
                try:
                    return float(cells[5])
                except ValueError:
                    return None
    return None


def host_info():
    """The host the CPU baselines run on: the CPUs this process may use, the
    count the OS reports, the model name and the OpenMP threads given."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0))
    return {"nproc": aff, "cpu_count": os.cpu_count(), "cpu_model": model,
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def run_original(nx: int, ny: int, ppc: int, steps: int, threads: int) -> float:
    """timePerStepMean of the compiled original's benchmark driver on a
    synthetic channel of nx x ny cells."""
    exe = os.path.join(ROOT, "oracle", "_ref", "main_benchmark_poiseuille_openmp_profiled_shifted_profiled")
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
        subprocess.run([exe, os.path.join(d, "params.kv"), os.path.join(d, "x.bin"), os.path.join(d, "v.bin"),
                        os.path.join(d, "t.bin"), os.path.join(d, "r0.bin"), os.path.join(d, "o"), "1", "1",
                        str(threads)], check=True, env=env, stdout=subprocess.DEVNULL, timeout=600)
        kv = {}
        with open(os.path.join(d, "o_runout.kv")) as fh:
            for line in fh:
                if "=" in line:
                    k, val = line.strip().split("=", 1)
                    kv[k] = val
        return float(kv["timePerStepMean"])
    finally:
        shutil.rmtree(d, ignore_errors=True)


def cpu_baseline(n_full: int, ppc: int, steps: int):
    """The compiled original (OpenMP, every thread this process is given) on
    two bounded samples, 128x64 and 256x128 cells.  Its step costs a*n +
    b*n^2 (each zone of a pass copies the whole state,
    zone_pass_openmp.cpp:67-72), so the two samples fix a and b and give the
    extrapolated time of one step of the full workload; ``value`` is the
    measured throughput of the larger sample."""
    exe = os.path.join(ROOT, "oracle", "_ref", "main_benchmark_poiseuille_openmp_profiled_shifted_profiled")
    if not os.path.exists(exe):
        return None
    host = host_info()
    threads = int(host["omp_num_threads_env"] or host["nproc"])
    samples = []
    for (nx, ny) in ((128, 64), (256, 128)):
        t = run_original(nx, ny, ppc, steps, threads)
        samples.append({"cells": [nx, ny], "particles": nx * ny * ppc, "s_per_step": t,
                        "particle_steps_per_s": nx * ny * ppc / t})
    (n1, t1), (n2, t2) = [(q["particles"], q["s_per_step"]) for q in samples]
    det = n1 * n2 * n2 - n2 * n1 * n1
    a = (t1 * n2 * n2 - t2 * n1 * n1) / det
    b = (n1 * t2 - n2 * t1) / det
    t_full = a * n_full + b * n_full * n_full
    return {
        "value": samples[1]["particle_steps_per_s"], "unit": "particle-steps/s", "cores": threads,
        "kind": "reference",
        "sample": (f"compiled original (main_benchmark_poiseuille_openmp_profiled_shifted_profiled, g++ -O3 -fopenmp, "
                   f"{threads} OpenMP threads) on the same physics at 256x128 cells x {ppc} ppc (n={n2}), {steps} "
                   f"steps; also 128x64 (n={n1}); value = the 256x128 sample's throughput"),
        "samples": samples, "host": host,
        "extrapolated_full_size": {"particles": n_full, "model": "t = a n + b n^2 fitted to the two samples",
                                   "a": a, "b": b, "s_per_step": t_full, "particle_steps_per_s": n_full / t_full},
    }


def cpu_port_full_size(p):
    """One step of the oracle's tile-local restatement (C++, one thread) on
    the full workload: a CPU baseline at the GPU's own size."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from tests.oracle_lib import CTR, get_oracle

    x, v, _, _ = uniform_state(p.n, p.Lx, p.Ly, seed=1)
    o = get_oracle()
    t0 = time.perf_counter()
    o.step(CTR, p, x, v, 1)
    dt = time.perf_counter() - t0
    return {"value": p.n / dt, "unit": "particle-steps/s", "cores": 1, "kind": "port", "s_per_step": dt,
            "sample": f"one step of the full {p.Nx}x{p.Ny} workload by the oracle's tile-local C++ restatement "
                      f"(oracle/, g++ -O2, one thread; O(n) per zone pass where the original is O(n^2))"}


def collect_samples(p_dict, forcings, per_forcing, warm, seed=1):
    """Closure samples (features, the plain closure's applied correction)
    from plain runs of the workload's grid, one run per body force, every
    second step after `warm` steps."""
    feats, tgts = [], []
    for gx in forcings:
        prm = dict(p_dict, bodyForceX=float(gx))
        ps = to_struct(prm)
        x, v, _, _ = uniform_state(ps.n, ps.Lx, ps.Ly, seed=seed)
        sim = gpu.Sim(ps)
        sim.set_state(x, v)
        del x, v
        sim.step(1, warm)
        sim.closure_sample(True, 1.0)
        s = warm + 1
        for _ in range(per_forcing):
            sim.step(s, 2)
            s += 2
            f, t = sim.closure_sample(True, 1.0, fetch=True)
            feats.append(f)
            tgts.append(t)
        sim.close()
    return np.stack(feats), np.stack(tgts)


def parabola_fit(prof, Ly):
    """Least-squares parabola plus slip, U(y) = a + b y (Ly - y), at the row
    centres."""
    Ny = prof.size
    y = (np.arange(Ny) + 0.5) * (Ly / Ny)
    A = np.stack([np.ones(Ny), y * (Ly - y)], axis=1)
    coef, *_ = np.linalg.lstsq(A, prof, rcond=None)
    return A @ coef


def poiseuille_profiles(p_dict, theta, out_scale, K, steps, runs):
    """The time- and x-averaged Ux(y) of the second half of `steps` steps of
    a channel, plain and with the surrogate (parameters theta), from the
    same `runs` initial states; summed on the device after every step."""
    from src_incompressible_openmp_amd import surrogate

    ps = to_struct(p_dict)
    net = surrogate.Net(ps.Nx, ps.Ny, True, False, max_batch=1)
    net.set_params(theta)
    out = {"plain": [], "learned": [], "none": []}
    counts = {"learned": 0, "plain": 0, "fallback": 0}
    half = steps // 2
    for r in range(runs):
        x, v, _, _ = uniform_state(ps.n, ps.Lx, ps.Ly, seed=100 + r)
        for mode in ("plain", "learned", "none"):
            sim = gpu.Sim(ps)
            sim.set_state(x, v)
            if mode == "learned":
                sim.set_surrogate(net, K, out_scale, 50.0 * out_scale)
            elif mode == "none":  # no closure correction on the surrogate's steps: the method without the network
                sim.set_surrogate(net, K, 1e-300, 1.0)
                sim.set_surrogate_check(1e300, 1e300)
            sim.step(1, half)
            sim.row_profile(1)
            sim.step(half + 1, steps - half)
            ux, _, cnt, _ = sim.row_profile(0, fetch=True)
            out[mode].append(ux / np.maximum(cnt, 1.0))
            if mode == "learned":
                c = sim.surrogate_counts()
                for k in counts:
                    counts[k] += c[k]
            sim.close()
    net.close()
    plain = np.mean(out["plain"], axis=0)
    learned = np.mean(out["learned"], axis=0)
    none = np.mean(out["none"], axis=0)
    fit = parabola_fit(plain, ps.Ly)
    nrm = np.linalg.norm(fit)
    return {"grid": [ps.Nx, ps.Ny], "particles": ps.n, "steps": steps, "window": [half + 1, steps], "runs": runs,
            "profile_err_plain": float(np.linalg.norm(plain - fit) / nrm),
            "profile_err_learned": float(np.linalg.norm(learned - fit) / nrm),
            "profile_err_no_correction": float(np.linalg.norm(none - fit) / nrm),
            "qx_plain": float(plain.mean() * ps.Ly), "qx_learned": float(learned.mean() * ps.Ly),
            "qx_no_correction": float(none.mean() * ps.Ly),
            "qx_rel_diff": float(abs(learned.mean() - plain.mean()) / max(abs(plain.mean()), 1e-300)),
            "profile_diff_learned_vs_plain": float(np.linalg.norm(learned - plain) / np.linalg.norm(plain)),
            "profile_diff_no_correction_vs_plain": float(np.linalg.norm(none - plain) / np.linalg.norm(plain)),
            "surrogate_steps": counts, "ux_plain": plain.tolist(), "ux_learned": learned.tolist(),
            "ux_no_correction": none.tolist()}


def learned_closure(sim, p, prm, s0, args, plain_value):
    """Trains the closure surrogate on the GPU from plain closures of the
    workload's grid at several body forces (1e-4 .. 1e-2), then times the
    workload with the surrogate (plain closure every K-th step and on a
    failed residual check, decided on the device), and measures the
    Poiseuille profile error of plain and surrogate runs on a 50x50
    channel of the same physics (where the flow develops within the run)."""
    from src_incompressible_openmp_amd import surrogate

    forcings = (1e-4, 1e-3, 1e-2)
    log(f"bench: closure surrogate: {args.train_samples} samples per body force {forcings}")
    t0 = time.perf_counter()
    feats, tgts = collect_samples(prm, forcings, args.train_samples, warm=20)
    t_sample = time.perf_counter() - t0
    scale = float(np.std(tgts)) or 1.0
    tgts /= scale
    rng = np.random.default_rng(1)
    theta = []
    for co, ci in ((32, 4), (32, 32), (32, 32), (2, 32)):
        b = np.sqrt(6.0 / (9 * ci))
        theta += [rng.uniform(-b, b, co * 9 * ci), np.zeros(co)]
    held = np.arange(len(feats)) % 8 == 7  # every 8th sample held out
    net = surrogate.Net(p.Nx, p.Ny, True, False, max_batch=4)
    net.set_params(np.concatenate(theta))
    log(f"bench: training {args.train_iters} Adam steps on {int((~held).sum())} samples")
    t0 = time.perf_counter()
    losses = net.train(feats[~held], tgts[~held], batch=4, iters=args.train_iters, lr=2e-3, seed=1)
    gpu.device_synchronize()
    t_train = time.perf_counter() - t0
    pred = net.forward(feats[held])
    r2 = float(1.0 - np.mean((pred - tgts[held]) ** 2) / np.mean(tgts[held] ** 2))
    theta_trained = net.get_params()
    del feats, tgts
    sim.set_surrogate(net, args.learned_K, scale, 50.0 * scale)
    sim.step(s0, args.warmup)
    s = s0 + args.warmup
    gpu.device_synchronize()
    log(f"bench: timing {args.steps} steps with the surrogate")
    c0 = sim.surrogate_counts()
    t0 = time.perf_counter()
    sim.step_async(s, args.steps)
    sim.synchronize()
    gpu.device_synchronize()
    el = time.perf_counter() - t0
    sim.check()
    c1 = sim.surrogate_counts()
    sim.set_surrogate(None, 0, 1.0, 1.0)
    net.close()
    value = p.n * args.steps / el
    log("bench: Poiseuille profile error, plain vs surrogate, 50x50 channel")
    prof = poiseuille_profiles(dict(prm, Nx=50, Ny=50, Nc=2500, n=50 * 50 * int(prm["gamma"]), Lx=10.0, Ly=10.0),
                               theta_trained, scale, args.learned_K, args.profile_steps, args.profile_runs)
    return {"_theta": theta_trained, "value": value, "ms_per_step": 1e3 * el / args.steps,
            "speedup_vs_plain": value / plain_value,
            "K": args.learned_K, "steps_learned": c1["learned"] - c0["learned"],
            "steps_plain": c1["plain"] - c0["plain"], "steps_fallback": c1["fallback"] - c0["fallback"],
            "train_forcings": list(forcings), "train_samples": int((~held).sum()), "heldout_samples": int(held.sum()),
            "train_iters": args.train_iters, "sample_s": t_sample, "train_s": t_train,
            "train_loss_first": float(np.mean(losses[:10])), "train_loss_last": float(np.mean(losses[-10:])),
            "heldout_r2": r2, "out_scale": scale, "profile": prof}


def large_instance(args, theta, out_scale):
    """The large single instance, 4096 x 2048 cells x 20 per cell (167.8M
    particles): plain steps, then (with trained weights) surrogate steps
    with the plain closure every K-th step; the 3x3 convolutions' weights do
    not depend on the grid, so the 1024 x 512 training carries over."""
    from src_incompressible_openmp_amd import surrogate

    Nx, Ny, ppc = 4096, 2048, 20
    prm = poiseuille_params(Nx=Nx, Ny=Ny, gamma=float(ppc))
    p = to_struct(prm)
    log(f"bench: large instance {Nx}x{Ny} x {ppc} (n={p.n}); building state")
    x, v, _, _ = uniform_state(p.n, p.Lx, p.Ly, seed=1)
    sim = gpu.Sim(p)
    sim.set_state(x, v)
    del x, v
    sim.step(1, 2)
    gpu.device_synchronize()
    steps = args.large_steps
    log(f"bench: large instance, timing {steps} plain steps")
    t0 = time.perf_counter()
    sim.step_async(3, steps)
    sim.synchronize()
    el = time.perf_counter() - t0
    sim.check()
    plain = p.n * steps / el
    out = {"workload": f"poiseuille_{Nx}x{Ny}_ppc{ppc}", "particles": p.n, "steps": steps,
           "plain": {"value": plain, "unit": "particle-steps/s", "ms_per_step": 1e3 * el / steps,
                     "step_roofline_frac": plain * STEP_ALG_BYTES / (HBM_PEAK_GBS * 1e9)}}
    if theta is not None:
        net = surrogate.Net(Nx, Ny, True, False, max_batch=1)
        net.set_params(theta)
        sim.set_surrogate(net, args.learned_K, out_scale, 50.0 * out_scale)
        s = 3 + steps
        sim.step(s, 2)
        s += 2
        gpu.device_synchronize()
        log(f"bench: large instance, timing {steps} surrogate steps")
        c0 = sim.surrogate_counts()
        t0 = time.perf_counter()
        sim.step_async(s, steps)
        sim.synchronize()
        el = time.perf_counter() - t0
        sim.check()
        c1 = sim.surrogate_counts()
        sim.set_surrogate(None, 0, 1.0, 1.0)
        net.close()
        val = p.n * steps / el
        out["learned"] = {"value": val, "unit": "particle-steps/s", "ms_per_step": 1e3 * el / steps,
                          "step_roofline_frac": val * STEP_ALG_BYTES / (HBM_PEAK_GBS * 1e9),
                          "speedup_vs_plain": val / plain, "K": args.learned_K,
                          "steps_learned": c1["learned"] - c0["learned"],
                          "steps_fallback": c1["fallback"] - c0["fallback"]}
    sim.close()
    return out


def ensemble_params(R: int, seed: int = 2):
    """R replicas of the Poiseuille channel: Nx uniform in 64..512, Ny in
    32..256, cell size 0.2, particles per cell in 5..20, body force
    log-uniform in [1e-4, 1e-2]; seed key r + 1 for replica r."""
    rng = np.random.default_rng(seed)
    prms = []
    for _ in range(R):
        nx, ny = int(rng.integers(64, 513)), int(rng.integers(32, 257))
        ppc = int(rng.integers(5, 21))
        gx = float(10.0 ** rng.uniform(-4.0, -2.0))
        prms.append(poiseuille_params(Nx=nx, Ny=ny, gamma=float(ppc), bodyForceX=gx))
    return prms, [r + 1 for r in range(R)]


def ensemble_leg(args):
    """The batched ensemble: args.ensemble_replicas uneven channels, each
    stage one launch over all replicas, the step replayed as a HIP graph;
    eager launches timed beside it (each kick but a call's last streams the
    next step).  Flow rates: each replica's Qx over as many further steps,
    its mean and the relative standard error of the mean (the channels
    start from rest, so the means follow the developing flow)."""
    R = args.ensemble_replicas
    prms, keys = ensemble_params(R)
    ps = [to_struct(q) for q in prms]
    n = sum(p.n for p in ps)
    log(f"bench: ensemble of {R} replicas, {n} particles; building states")
    ens = gpu.Ensemble(ps, keys)
    for r, p in enumerate(ps):
        x, v, _, _ = uniform_state(p.n, p.Lx, p.Ly, seed=1000 + r)
        ens.set_state(r, x, v)
    del x, v
    ens.step(1, 2, graph=True)
    steps = args.ensemble_steps
    log(f"bench: ensemble, timing {steps} graph-replayed steps")
    t0 = time.perf_counter()
    ens.step(3, steps, graph=True, wait=False)
    ens.synchronize()
    el = time.perf_counter() - t0
    k = max(3, steps // 4)
    log(f"bench: ensemble, timing {k} eagerly launched steps")
    t0 = time.perf_counter()
    ens.step(3 + steps, k, graph=False, wait=False)
    ens.synchronize()
    el_eager = time.perf_counter() - t0
    log(f"bench: ensemble, flow rates over {steps} more steps")
    ens.flow(True)
    ens.step(3 + steps + k, steps, graph=True)
    q, q2, ns = ens.flow(False, fetch=True)
    info = ens.info()
    ens.close()
    mean = q / ns
    var = np.maximum(q2 / ns - mean * mean, 0.0)
    rel_se = np.sqrt(var / max(ns - 1, 1)) / np.maximum(np.abs(mean), 1e-300)
    return {"workload": f"ensemble_{R}_poiseuille_uneven", "replicas": R, "particles": n, "cells": info["cells"],
            "zones": info["zones"], "steps": steps, "value": n * steps / el, "unit": "particle-steps/s",
            "ms_per_step": 1e3 * el / steps, "step_roofline_frac": n * steps / el * STEP_ALG_BYTES / (HBM_PEAK_GBS * 1e9),
            "eager_ms_per_step": 1e3 * el_eager / k, "graph_speedup": (el_eager / k) / (el / steps),
            "flow": {"qx_mean_min": float(mean.min()), "qx_mean_max": float(mean.max()),
                     "ensemble_mean_rel_stderr": float(rel_se.mean()), "steps": ns},
            "grid_ranges": {"Nx": [min(p.Nx for p in ps), max(p.Nx for p in ps)],
                            "Ny": [min(p.Ny for p in ps), max(p.Ny for p in ps)],
                            "ppc": [min(int(p.gamma) for p in ps), max(int(p.gamma) for p in ps)]}}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--ny", type=int, default=512)
    ap.add_argument("--ppc", type=int, default=20)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cpu-port", action="store_true", help="skip the one-step oracle timing at full size")
    ap.add_argument("--no-profile", action="store_true", help="time without per-kernel events")
    ap.add_argument("--no-learned", action="store_true", help="skip the closure-surrogate measurement")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 83,886,080 bytes device to device 5 times (the counter calibration of the "
                         "traffic figures) and exit")
    ap.add_argument("--learned-K", type=int, default=50, help="plain closure every K-th step with the surrogate")
    ap.add_argument("--train-samples", type=int, default=40, help="closure samples per body force")
    ap.add_argument("--train-iters", type=int, default=600)
    ap.add_argument("--profile-steps", type=int, default=3000)
    ap.add_argument("--profile-runs", type=int, default=8)
    ap.add_argument("--no-large", action="store_true", help="skip the 4096x2048 large-instance measurement")
    ap.add_argument("--large-steps", type=int, default=10)
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--ensemble-replicas", type=int, default=256)
    ap.add_argument("--ensemble-steps", type=int, default=20)
    ap.add_argument("--ensemble-only", action="store_true",
                    help="only the batched-ensemble measurement (its JSON object; for profiling)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch

        a = torch.ones(10485760, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log("bench: copy calibration done (5 x 83,886,080 bytes)")
        return 0

    if args.ensemble_only:
        print(json.dumps({"ensemble": ensemble_leg(args)}), flush=True)
        return 0

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the decomposition into x-slabs over several GPUs is designed, not implemented)", file=sys.stderr)
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
    log(f"bench: timing {args.steps} steps")
    t0 = time.perf_counter()
    sim.step_async(1 + args.warmup, args.steps)
    sim.synchronize()
    gpu.device_synchronize()
    t1 = time.perf_counter()
    sim.check()
    elapsed = t1 - t0
    s_next = 1 + args.warmup + args.steps
    prof = None
    profiled_ms = None
    if not args.no_profile:
        # the per-kernel HIP events (kernel times, roofline) on as many more
        # steps: an event between kernels costs the stream a gap, so the
        # throughput above is timed without them
        log(f"bench: {args.steps} more steps with per-kernel HIP events")
        sim.profile(True)
        t2 = time.perf_counter()
        sim.step_async(s_next, args.steps)
        sim.synchronize()
        gpu.device_synchronize()
        profiled_ms = 1e3 * (time.perf_counter() - t2) / args.steps
        sim.check()
        prof = sim.profile_read()
        sim.profile(False)
        s_next += args.steps
    m = sim.step(s_next, 1, metrics=True)
    s_next += 1
    value = n * args.steps / elapsed
    learned = None if args.no_learned else learned_closure(sim, p, prm, s_next, args, value)
    sim.close()
    theta = learned.pop("_theta") if learned else None
    large = None if args.no_large else large_instance(args, theta, learned["out_scale"] if learned else 1.0)
    ensemble = None if args.no_ensemble else ensemble_leg(args)

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
        out["profiled_ms_per_step"] = profiled_ms  # the steps the events timed, gaps included
        cands = {k: per[k]["ms_total"] for k in ALG_BYTES if per[k]["launches"]}
        dom = max(cands, key=cands.get)
        dur = per[dom]["ms_per_launch"] * 1e-3
        alg = ALG_BYTES[dom] * n
        ach = alg / dur / 1e9
        traffic = pmc_traffic(KERNEL_NAME[dom], out["config"]["workload"])
        out["roofline"] = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                           "alg_bytes_per_launch": alg, "ms_per_launch": per[dom]["ms_per_launch"]}
        # every particle kernel class against the same bound, for comparison
        out["kernel_roofline"] = {
            k: {"ms_per_launch": per[k]["ms_per_launch"],
                "achieved_GBs": ALG_BYTES[k] * n / (per[k]["ms_per_launch"] * 1e-3) / 1e9,
                "frac": ALG_BYTES[k] * n / (per[k]["ms_per_launch"] * 1e-3) / 1e9 / HBM_PEAK_GBS}
            for k in ALG_BYTES if per[k]["launches"]}
    if learned:
        out["learned"] = learned
    if large:
        out["large_instance"] = large
    if ensemble:
        out["ensemble"] = ensemble
    out["last_step"] = {"moved_dense": m["base"]["nParticlesMovedDense"] + m["shifted"]["nParticlesMovedDense"],
                        "moved_sparse": m["base"]["nParticlesMovedSparse"] + m["shifted"]["nParticlesMovedSparse"],
                        "beta_applied": m["closure"]["betaRepairApplied"]}
    if not args.no_cpu_baseline:
        log(f"bench: CPU baseline (compiled original, 128x64 and 256x128 x {args.ppc}, {args.cpu_steps} steps each)")
        cb = cpu_baseline(n, args.ppc, args.cpu_steps)
        out["cpu_baseline"] = cb
        if cb:
            out["gpu_over_cpu_extrapolated_full_size"] = value / cb["extrapolated_full_size"]["particle_steps_per_s"]
        if not args.no_cpu_port:
            log("bench: CPU port, one step of the full workload")
            cp = cpu_port_full_size(p)
            out["cpu_port_full_size"] = cp
            out["gpu_over_cpu_port_full_size"] = value / cp["value"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
