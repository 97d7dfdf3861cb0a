"""Generate the golden fixtures from the compiled original reference.

Runs the reference's own drivers (built from its sources by oracle/Makefile
into oracle/_ref/) on the cases of cases.py and stores, per case, the inputs
and the outputs as plain arrays in ``<case>.npz`` (no pickles), plus SHA-256
digests of the outputs of the larger cases in ``digests.json``.

Run in the development container, where the reference sources exist:

    make -C oracle && python tests/golden/make_goldens.py
    python tests/golden/make_goldens.py --zone-kernel   # zonekernel.npz only
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from src_incompressible_openmp_amd import fileio  # noqa: E402
from src_incompressible_openmp_amd.params import write_params_kv  # noqa: E402
from tests.golden.cases import DIGEST_CASES, SMALL_CASES, case_params, case_state, ref_red_pair  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")


def run(exe: str, *args: str) -> None:
    subprocess.run([os.path.join(REF, exe), *args], check=True, env=dict(os.environ, OMP_NUM_THREADS="2"),
                   stdout=subprocess.DEVNULL)


def read_kv(path: str):
    out = {}
    with open(path) as fh:
        for line in fh:
            line = line.strip()
            if "=" in line:
                k, val = line.split("=", 1)
                try:
                    out[k] = float(val)
                except ValueError:
                    pass
    return out


def write_inputs(d: str, p, x, v, t, r0, extra=None):
    write_params_kv(os.path.join(d, "params.kv"), p, extra)
    n = int(p["n"])
    fileio.write_xy(os.path.join(d, "x.bin"), x, n)
    fileio.write_xy(os.path.join(d, "v.bin"), v, n)
    fileio.write_u8(os.path.join(d, "type.bin"), t, n)
    fileio.write_xy(os.path.join(d, "r0.bin"), r0, n)


def kv_arrays(prefix: str, kv):
    keys = sorted(kv)
    return {prefix + "_kv_keys": np.array(keys), prefix + "_kv_vals": np.array([kv[k] for k in keys])}


def cf_arrays(prefix: str, path: str):
    cf = fileio.read_cell_fields(path)
    out = {prefix + "_N": cf["N"]}
    for k in fileio.CELL_FIELD_NAMES:
        out[prefix + "_" + k] = cf[k]
    return out


def small_case(name: str):
    p = case_params(name)
    n = int(p["n"])
    x, v, t, r0 = case_state(name)
    out = {"x0": x, "v0": v}
    bench_steps = 6
    with tempfile.TemporaryDirectory() as d:
        write_inputs(d, p, x, v, t, r0, {"benchmark_nSteps": bench_steps, "benchmark_metricsEvery": 1,
                                         "benchmark_dumpSteps": str(bench_steps)})
        ins = [os.path.join(d, f) for f in ("params.kv", "x.bin", "v.bin", "type.bin", "r0.bin")]
        o = os.path.join(d, "o")
        # tranche 1: force + stream + BC + collision (seeds 104729, 130363, 433494437)
        run("main_step01_stream_bc_collision", *ins, o + "1", "1", "1")
        out["s1_x"] = fileio.read_xy(o + "1_x.bin", n)
        out["s1_v"] = fileio.read_xy(o + "1_v.bin", n)
        out.update(cf_arrays("s1_cf", o + "1_cellfields.bin"))
        out.update(kv_arrays("s1", read_kv(o + "1_runout.kv")))
        # tranche 2: one Gauss-Seidel zone pass per layout on the input state
        for lay in ("base", "shifted"):
            run("main_step02_zone_pass", *ins, o + "2" + lay, lay, "1", "1")
            out[f"s2{lay}_x"] = fileio.read_xy(o + "2" + lay + "_x.bin", n)
            out[f"s2{lay}_v"] = fileio.read_xy(o + "2" + lay + "_v.bin", n)
            out[f"s2{lay}_owner"] = fileio.read_grid(o + "2" + lay + "_zone_owner.bin", "<i4")[2]
            out.update(cf_arrays(f"s2{lay}_before", o + "2" + lay + "_cellfields_before.bin"))
            out.update(cf_arrays(f"s2{lay}_after", o + "2" + lay + "_cellfields_after.bin"))
            out.update(kv_arrays(f"s2{lay}", read_kv(o + "2" + lay + "_runout.kv")))
        # tranche 4: closure on an ad-hoc (ref, red) pair
        xr, vr = ref_red_pair(x, v, float(p["Lx"]), float(p["Ly"]))
        out["s4_xred"], out["s4_vred"] = xr, vr
        fileio.write_xy(os.path.join(d, "xr.bin"), xr, n)
        fileio.write_xy(os.path.join(d, "vr.bin"), vr, n)
        run("main_step04_liquid_closure", ins[0], ins[1], ins[2], os.path.join(d, "xr.bin"),
            os.path.join(d, "vr.bin"), ins[3], ins[4], o + "4", "1", "1", "1")
        out["s4_x"] = fileio.read_xy(o + "4_x.bin", n)
        out["s4_v"] = fileio.read_xy(o + "4_v.bin", n)
        for f in ("ref", "red", "repaired", "out"):
            out.update(cf_arrays("s4_cf" + f, o + f"4_cellfields_{f}.bin"))
        for g in ("pdrive_raw", "pdrive", "durepairx", "durepairy", "dueosx", "dueosy"):
            out["s4_" + g] = fileio.read_grid(o + f"4_{g}.bin", "<f8")[2]
        out["s4_mask"] = fileio.read_grid(o + "4_interzone_mask.bin", "<i4")[2]
        out.update(kv_arrays("s4", read_kv(o + "4_runout.kv")))
        # tranche 5: full step with Gauss-Seidel passes
        run("main_step05_full_step", *ins, o + "5", "1", "1")
        for stg in ("ref", "after_base", "red"):
            out[f"s5_{stg}_x"] = fileio.read_xy(o + f"5_{stg}_x.bin", n)
            out[f"s5_{stg}_v"] = fileio.read_xy(o + f"5_{stg}_v.bin", n)
        out["s5_x"] = fileio.read_xy(o + "5_x.bin", n)
        out["s5_v"] = fileio.read_xy(o + "5_v.bin", n)
        out.update(kv_arrays("s5", read_kv(o + "5_runout.kv")))
        # step 9: snapshot passes (serial and OpenMP) + closure
        run("main_step09_openmp_full_validate_syncserial", *ins, o + "9", "1", "1", "2")
        for stg in ("ref", "serial_after_base", "serial", "omp_after_base", "omp"):
            out[f"s9_{stg}_x"] = fileio.read_xy(o + f"9_{stg}_x.bin", n)
            out[f"s9_{stg}_v"] = fileio.read_xy(o + f"9_{stg}_v.bin", n)
        out.update(kv_arrays("s9", read_kv(o + "9_runout.kv")))
        # benchmark driver: steps 1..6 with snapshot passes
        run("main_benchmark_poiseuille_openmp_profiled_shifted_profiled", *ins, o + "b", "1", "1", "2")
        stem = o + f"b_step{bench_steps:04d}"
        out["bench_x"] = fileio.read_xy(stem + "_x.bin", n)
        out["bench_v"] = fileio.read_xy(stem + "_v.bin", n)
        out.update(cf_arrays("bench_cf", stem + "_cellfields.bin"))
        with open(o + "b_metrics.csv") as fh:
            rows = [ln.strip().split(",") for ln in fh if ln.strip()]
        out["bench_metrics_header"] = np.array(rows[0])
        out["bench_metrics"] = np.array([[float(c) for c in r] for r in rows[1:]])
        out["bench_steps"] = np.array(bench_steps)
    return out


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<f8").tobytes()).hexdigest()


def digest_case(name: str):
    p = case_params(name)
    n = int(p["n"])
    x, v, t, r0 = case_state(name)
    res = {"x0": digest(x), "v0": digest(v)}
    steps = 3
    with tempfile.TemporaryDirectory() as d:
        write_inputs(d, p, x, v, t, r0, {"benchmark_nSteps": steps, "benchmark_metricsEvery": 1,
                                         "benchmark_dumpSteps": str(steps)})
        ins = [os.path.join(d, f) for f in ("params.kv", "x.bin", "v.bin", "type.bin", "r0.bin")]
        o = os.path.join(d, "o")
        run("main_step09_openmp_full_validate_syncserial", *ins, o + "9", "1", "1", "4")
        for stg in ("ref", "serial_after_base", "serial"):
            res[f"s9_{stg}_x"] = digest(fileio.read_xy(o + f"9_{stg}_x.bin", n))
            res[f"s9_{stg}_v"] = digest(fileio.read_xy(o + f"9_{stg}_v.bin", n))
        run("main_benchmark_poiseuille_openmp_profiled_shifted_profiled", *ins, o + "b", "1", "1", "4")
        stem = o + f"b_step{steps:04d}"
        res["bench_steps"] = steps
        res["bench_x"] = digest(fileio.read_xy(stem + "_x.bin", n))
        res["bench_v"] = digest(fileio.read_xy(stem + "_v.bin", n))
    return res


ZONE_KERNEL_CASES = SMALL_CASES + DIGEST_CASES


def zone_kernel_ids(n_zones: int):
    """zones probed per layout: the first, one in the middle, the last"""
    return sorted({0, n_zones // 2, n_zones - 1})


def zone_kernel_case(name: str):
    """run_zone_kernel_inplace of the original (through oracle/_ref/zone_kernel_probe)
    on single zones of both layouts, from the case's input state: its
    ZoneKernelStats (stats, diag, moved counts) and SHA-256 digests of the
    state it leaves."""
    p = case_params(name)
    n = int(p["n"])
    x, v, _, _ = case_state(name)
    out = {}
    with tempfile.TemporaryDirectory() as d:
        write_params_kv(os.path.join(d, "params.kv"), p)
        fileio.write_xy(os.path.join(d, "x.bin"), x, n)
        fileio.write_xy(os.path.join(d, "v.bin"), v, n)
        for lay in ("base", "shifted"):
            zones, zi = None, 0
            todo = [0]
            while todo:
                zi = todo.pop(0)
                seed = 7001 + zi
                run("zone_kernel_probe", os.path.join(d, "params.kv"), os.path.join(d, "x.bin"),
                    os.path.join(d, "v.bin"), lay, str(zi), str(seed), os.path.join(d, "o"))
                kv = read_kv(os.path.join(d, "o_stats.kv"))
                if zones is None:
                    zones = int(kv["nZones"])
                    todo = [z for z in zone_kernel_ids(zones) if z != 0]
                key = f"{name}_{lay}_{zi}"
                out[key + "_seed"] = np.array(seed)
                out[key + "_stats"] = np.array([kv[f"stats{k}"] for k in range(6)])
                out[key + "_diag"] = np.array([kv[f"diag{k}"] for k in range(10)])
                out[key + "_moved"] = np.array([kv["movedDense"], kv["movedSparse"]], np.int32)
                out[key + "_x_sha256"] = np.array(digest(fileio.read_xy(os.path.join(d, "o_x.bin"), n)))
                out[key + "_v_sha256"] = np.array(digest(fileio.read_xy(os.path.join(d, "o_v.bin"), n)))
            out[f"{name}_{lay}_zones"] = np.array(zones)
    return out


MASK_ZONE_CASES = ("small", "override", "cfg0")


def zone_masks(Nx: int, Ny: int):
    """Zones that are not rectangles, as ZoneDescriptor.maskVec (one byte per
    cell, c = iy + Ny*ix): a rectangle with a hole, an L, a random 40% of the
    interior, and a band across the periodic x boundary."""
    def grid():
        return np.zeros((Nx, Ny), np.uint8)
    notch = grid()
    notch[1:min(Nx, 9), 1:min(Ny, 8)] = 1
    notch[3:5, 3:5] = 0
    ell = grid()
    ell[: Nx // 2, :3] = 1
    ell[:3, : (2 * Ny) // 3] = 1
    rng = np.random.default_rng(17 + Nx)
    scatter = grid()
    scatter[2:Nx - 2, 1:Ny - 1] = (rng.random((Nx - 4, Ny - 2)) < 0.4).astype(np.uint8)
    wrap = grid()
    wrap[[Nx - 2, Nx - 1, 0, 1], 2:6] = 1
    return {k: m.reshape(-1) for k, m in (("notch", notch), ("ell", ell), ("scatter", scatter), ("wrap", wrap))}


def zone_kernel_mask_case(name: str):
    """run_zone_kernel_inplace of the original on the zones of zone_masks
    (through zone_kernel_probe's mask argument): stats, diag, moved counts
    and SHA-256 digests of the state it leaves, with each mask."""
    p = case_params(name)
    n = int(p["n"])
    x, v, _, _ = case_state(name)
    out = {}
    with tempfile.TemporaryDirectory() as d:
        write_params_kv(os.path.join(d, "params.kv"), p)
        fileio.write_xy(os.path.join(d, "x.bin"), x, n)
        fileio.write_xy(os.path.join(d, "v.bin"), v, n)
        for mi, (mname, mask) in enumerate(zone_masks(int(p["Nx"]), int(p["Ny"])).items()):
            seed = 9001 + mi
            mask.astype(np.uint8).tofile(os.path.join(d, "mask.bin"))
            run("zone_kernel_probe", os.path.join(d, "params.kv"), os.path.join(d, "x.bin"),
                os.path.join(d, "v.bin"), "base", "0", str(seed), os.path.join(d, "o"), os.path.join(d, "mask.bin"))
            kv = read_kv(os.path.join(d, "o_stats.kv"))
            key = f"{name}_mask_{mname}"
            out[key] = mask.astype(np.uint8)
            out[key + "_seed"] = np.array(seed)
            out[key + "_stats"] = np.array([kv[f"stats{k}"] for k in range(6)])
            out[key + "_diag"] = np.array([kv[f"diag{k}"] for k in range(10)])
            out[key + "_moved"] = np.array([kv["movedDense"], kv["movedSparse"]], np.int32)
            out[key + "_x_sha256"] = np.array(digest(fileio.read_xy(os.path.join(d, "o_x.bin"), n)))
            out[key + "_v_sha256"] = np.array(digest(fileio.read_xy(os.path.join(d, "o_v.bin"), n)))
    return out


def main() -> None:
    if not os.path.exists(os.path.join(REF, "main_step01_stream_bc_collision")):
        raise SystemExit("build the compiled original first: make -C oracle")
    if "--zone-kernel" in sys.argv[1:]:
        arrs = {}
        for name in ZONE_KERNEL_CASES:
            arrs.update(zone_kernel_case(name))
        for name in MASK_ZONE_CASES:
            arrs.update(zone_kernel_mask_case(name))
        np.savez_compressed(os.path.join(HERE, "zonekernel.npz"), **arrs)
        print("wrote zonekernel.npz")
        return
    for name in SMALL_CASES:
        arrs = small_case(name)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **arrs)
        print("wrote", name, sum(a.nbytes for a in arrs.values()) // 1024, "KiB raw")
    dig = {name: digest_case(name) for name in DIGEST_CASES}
    with open(os.path.join(HERE, "digests.json"), "w") as fh:
        json.dump(dig, fh, indent=1, sort_keys=True)
    print("wrote digests.json")


if __name__ == "__main__":
    main()
