"""The C-ABI library loads on a host without a GPU and exports every entry
point the public headers declare (no compute calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = []
    for fn in sorted(os.listdir(INCLUDE)):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        inside = text.split('extern "C" {', 1)
        if len(inside) < 2:
            continue  # header-only (inline) code, nothing exported
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(mpcd_\w+)\s*\(", inside[1], flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    from src_incompressible_openmp_amd import build

    if not os.path.exists(build.LIB):
        build.build()
    return ctypes.CDLL(build.LIB)


def test_headers_declare_entry_points():
    names = declared_functions()
    assert "mpcd_gpu_sim_step" in names and "mpcd_gpu_zone_pass" in names
    assert len(names) >= 20


@pytest.mark.parametrize("name", declared_functions())
def test_symbol_exported(lib, name):
    assert hasattr(lib, name), f"{name} declared in include/ but not exported"


def test_error_path_without_gpu_is_an_error_code(lib):
    """With no visible device the library reports an error code and a message
    instead of crashing (a GPU box takes the normal path)."""
    lib.mpcd_gpu_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
    lib.mpcd_gpu_last_error.restype = ctypes.c_char_p
    n = ctypes.c_int(-1)
    rc = lib.mpcd_gpu_device_count(ctypes.byref(n))
    if rc == 0:
        assert n.value >= 0
    else:
        assert rc < 0 and lib.mpcd_gpu_last_error()


def test_no_particles_is_refused_before_any_device_call():
    """n = 0 is refused with MPCD_E_INVALID and a message by the simulation
    and by the stateless entry points that build one, before any device
    call (so the same holds on a box without a GPU)."""
    from src_incompressible_openmp_amd import gpu
    from src_incompressible_openmp_amd.params import to_struct
    from tests.golden.cases import case_params

    p = to_struct(dict(case_params("small"), n=0))
    with pytest.raises(gpu.MpcdError, match="no particles") as e:
        gpu.Sim(p)
    assert e.value.code == -1  # MPCD_E_INVALID
    with pytest.raises(gpu.MpcdError, match="no particles"):
        gpu.compute_cell_fields(p, np.zeros(0), np.zeros(0), 10.0)


def test_profile_classes_match_header():
    """The Python binding names one timing class per MPCD_PROF_* entry, in
    the header's order."""
    import re

    from src_incompressible_openmp_amd import gpu

    text = open(os.path.join(INCLUDE, "mpcd_gpu.h")).read()
    entries = dict((k, int(v)) for k, v in re.findall(r"MPCD_PROF_([A-Z_]+)\s*=\s*(\d+)", text))
    n = entries.pop("CLASSES")
    assert n == len(gpu.PROF_CLASSES)
    for name, idx in entries.items():
        assert gpu.PROF_CLASSES[idx] == name.lower(), (name, idx)
