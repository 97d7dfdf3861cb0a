"""The no-allocation entry point mpcd_gpu_compute_cell_fields_noalloc
(device buffers in and out, static scratch) against the allocating
mpcd_gpu_compute_cell_fields on the same inputs: every field bit for bit,
on two calls in a row, the second with a smaller input (the scratch keeps
the first call's contents), and the size limit rejected before any work."""
import ctypes

import numpy as np
import pytest

from src_incompressible_openmp_amd.params import poiseuille_params, to_struct
from src_incompressible_openmp_amd.fileio import uniform_state

pytestmark = pytest.mark.gpu

FIELDS = ("Ux", "Uy", "Px", "Py", "T", "rho", "Pkin", "Pvir", "P", "rhoTarget")


def hip_runtime(gpu):
    """The HIP runtime the library itself loaded (from the process's maps),
    so that device buffers come from the same runtime instance."""
    gpu.load()
    with open("/proc/self/maps") as fh:
        paths = {ln.split()[-1] for ln in fh if "libamdhip64" in ln and "/" in ln}
    assert paths, "HIP runtime not mapped"
    hip = ctypes.CDLL(sorted(paths)[0])
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


H2D, D2H = 1, 2


def dalloc(hip, nbytes):
    ptr = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(ptr), max(int(nbytes), 8)) == 0
    return ptr


def run_noalloc(gpu, hip, p, x, v, rt):
    lib = gpu.load()
    bufs = []
    xd, vd = dalloc(hip, x.nbytes), dalloc(hip, v.nbytes)
    bufs += [xd, vd]
    assert hip.hipMemcpy(xd, x.ctypes.data, x.nbytes, H2D) == 0
    assert hip.hipMemcpy(vd, v.ctypes.data, v.nbytes, H2D) == 0
    host = {"N": np.full(p.Nc, -7, np.int32)}
    for k in FIELDS:
        host[k] = np.full(p.Nc, np.nan)
    dev = {}
    for k, a in host.items():  # outputs start as garbage the call must overwrite
        dev[k] = dalloc(hip, a.nbytes)
        bufs.append(dev[k])
        assert hip.hipMemcpy(dev[k], a.ctypes.data, a.nbytes, H2D) == 0
    dp = ctypes.POINTER(ctypes.c_double)
    f = gpu.CellFieldsC(ctypes.cast(dev["N"], ctypes.POINTER(ctypes.c_int32)),
                        *[ctypes.cast(dev[k], dp) for k in FIELDS])
    rc = lib.mpcd_gpu_compute_cell_fields_noalloc(ctypes.byref(p), ctypes.cast(xd, dp), ctypes.cast(vd, dp),
                                                   ctypes.c_double(rt), ctypes.byref(f), None)
    gpu._check(rc)
    assert hip.hipDeviceSynchronize() == 0
    for k, a in host.items():
        assert hip.hipMemcpy(a.ctypes.data, dev[k], a.nbytes, D2H) == 0
    for b in bufs:
        hip.hipFree(b)
    return host


def test_noalloc_cell_fields_match_allocating_entry_point():
    from src_incompressible_openmp_amd import gpu

    gpu.load()
    assert gpu.device_count() >= 1
    hip = hip_runtime(gpu)
    for (nx, ny, ppc, seed) in ((96, 64, 20, 3), (40, 20, 7, 4)):  # second call smaller
        p = to_struct(poiseuille_params(Nx=nx, Ny=ny, gamma=float(ppc)))
        x, v, _, _ = uniform_state(p.n, p.Lx, p.Ly, seed=seed)
        rt = p.gamma / ((p.Lx / p.Nx) * (p.Ly / p.Ny))
        want = gpu.compute_cell_fields(p, x, v, rt)
        got = run_noalloc(gpu, hip, p, x, v, rt)
        for k in ("N",) + FIELDS:
            assert np.asarray(got[k]).tobytes() == np.asarray(want[k]).tobytes(), (nx, k)
    big = to_struct(poiseuille_params(Nx=1024, Ny=512, gamma=1.0))  # 524,288 cells: beyond the limit
    lib = gpu.load()
    f = gpu.CellFieldsC()
    dummy = ctypes.cast(ctypes.c_void_p(16), ctypes.POINTER(ctypes.c_double))
    rc = lib.mpcd_gpu_compute_cell_fields_noalloc(ctypes.byref(big), dummy, dummy, ctypes.c_double(1.0),
                                                   ctypes.byref(f), None)
    assert rc == -1 and b"at most" in lib.mpcd_gpu_last_error()
