"""Run parameters: the strict ``params.kv`` format and the C-ABI image.

``read_params_kv`` follows the reference reader (params_io.cpp:21-137): one
``key=value`` per line, ``#`` comments, surrounding whitespace trimmed, every
one of the 53 keys required, booleans ``true/false/1/0``.  ``write_params_kv``
writes numbers with ``%.17g`` as the reference's MATLAB writer does
(write_params_kv.m:17).  ``MpcdParams`` is the ctypes mirror of
``mpcd_params`` in include/mpcd_gpu.h.
"""
from __future__ import annotations

import ctypes
import math
from typing import Dict, Mapping, Tuple

# (key, kind) in the order of mpcd::Params (types.h:10-72)
PARAM_KEYS: Tuple[Tuple[str, str], ...] = (
    ("Nx", "int"), ("Ny", "int"), ("Nc", "int"), ("n", "int"),
    ("Lx", "float"), ("Ly", "float"), ("dt", "float"), ("a0", "float"),
    ("alpha", "float"), ("kBT", "float"), ("g", "float"), ("bodyForceX", "float"),
    ("useThermostat", "bool"), ("keepMeanFlow", "bool"),
    ("boundary_left", "str"), ("boundary_right", "str"), ("boundary_bottom", "str"), ("boundary_top", "str"),
    ("Utop", "float"), ("Ubottom", "float"), ("wallSigma", "float"),
    ("noSurfaceCase", "bool"), ("redistributionEnableSurfaceTopology", "bool"),
    ("redistributionWallWettingEnabled", "bool"), ("useInterfaceVelocityReorientation", "bool"),
    ("useIncompressibleRedistribution", "bool"), ("redistribAfterCollision", "bool"),
    ("useZoneRedistribution", "bool"), ("zoneTileNx", "int"), ("zoneTileNy", "int"),
    ("zoneUseShiftedSecondPass", "bool"), ("zonePassthroughMode", "bool"), ("zoneSingleFullDomainMode", "bool"),
    ("enableMomentumCorrectionPostRedistribution", "bool"), ("gamma", "float"), ("coef", "float"),
    ("highMode", "str"), ("lowMode", "str"), ("maxRedistribPasses", "int"),
    ("useLocalFluidFractionThresholds", "bool"), ("fluidFracGain", "float"), ("lowThrFloor", "float"),
    ("highThrFloor", "float"), ("lowThrBulkOverride", "float"),
    ("useLiquidClosure", "bool"), ("useOptimalBetaRepair", "bool"), ("betaRepair", "float"), ("betaEOS", "float"),
    ("Kvirial", "float"), ("smoothPdriveAtInterzone", "bool"), ("smoothPdriveInterzoneWidthCells", "int"),
    ("smoothPdriveInterzoneBlend", "float"), ("smoothPdriveIncludeShiftedLayouts", "bool"),
)

BC_MODES = {"periodic": 0, "thermalize": 1, "bounceback": 2, "specular": 3}

LAYOUT_BASE, LAYOUT_SHIFTED, LAYOUT_SINGLE_FULL = 0, 1, 2


class MpcdParams(ctypes.Structure):
    """ctypes image of ``mpcd_params`` (include/mpcd_gpu.h)."""

    _fields_ = [
        ("Nx", ctypes.c_int32), ("Ny", ctypes.c_int32), ("Nc", ctypes.c_int32), ("n", ctypes.c_int32),
        ("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("dt", ctypes.c_double), ("a0", ctypes.c_double),
        ("alpha", ctypes.c_double), ("kBT", ctypes.c_double), ("g", ctypes.c_double), ("bodyForceX", ctypes.c_double),
        ("useThermostat", ctypes.c_int32), ("keepMeanFlow", ctypes.c_int32),
        ("bc_left", ctypes.c_int32), ("bc_right", ctypes.c_int32), ("bc_bottom", ctypes.c_int32),
        ("bc_top", ctypes.c_int32),
        ("Utop", ctypes.c_double), ("Ubottom", ctypes.c_double), ("wallSigma", ctypes.c_double),
        ("noSurfaceCase", ctypes.c_int32), ("redistributionEnableSurfaceTopology", ctypes.c_int32),
        ("redistributionWallWettingEnabled", ctypes.c_int32), ("useInterfaceVelocityReorientation", ctypes.c_int32),
        ("useIncompressibleRedistribution", ctypes.c_int32), ("redistribAfterCollision", ctypes.c_int32),
        ("useZoneRedistribution", ctypes.c_int32),
        ("zoneTileNx", ctypes.c_int32), ("zoneTileNy", ctypes.c_int32),
        ("zoneUseShiftedSecondPass", ctypes.c_int32), ("zonePassthroughMode", ctypes.c_int32),
        ("zoneSingleFullDomainMode", ctypes.c_int32), ("enableMomentumCorrectionPostRedistribution", ctypes.c_int32),
        ("gamma", ctypes.c_double), ("coef", ctypes.c_double),
        ("high_mode_one_plus_coef", ctypes.c_int32), ("low_mode_one_plus_coef", ctypes.c_int32),
        ("maxRedistribPasses", ctypes.c_int32), ("useLocalFluidFractionThresholds", ctypes.c_int32),
        ("fluidFracGain", ctypes.c_double), ("lowThrFloor", ctypes.c_double), ("highThrFloor", ctypes.c_double),
        ("lowThrBulkOverride", ctypes.c_double),
        ("useLiquidClosure", ctypes.c_int32), ("useOptimalBetaRepair", ctypes.c_int32),
        ("betaRepair", ctypes.c_double), ("betaEOS", ctypes.c_double), ("Kvirial", ctypes.c_double),
        ("smoothPdriveAtInterzone", ctypes.c_int32), ("smoothPdriveInterzoneWidthCells", ctypes.c_int32),
        ("smoothPdriveInterzoneBlend", ctypes.c_double),
        ("smoothPdriveIncludeShiftedLayouts", ctypes.c_int32), ("_pad", ctypes.c_int32),
    ]


def _parse_bool(s: str) -> bool:
    if s in ("true", "1"):
        return True
    if s in ("false", "0"):
        return False
    raise ValueError(f"Invalid boolean value: {s}")


def _parse_int(s: str) -> int:
    # std::stoi semantics: leading integer prefix, error if none
    s2 = s.strip()
    j = 0
    if j < len(s2) and s2[j] in "+-":
        j += 1
    k = j
    while k < len(s2) and s2[k].isdigit():
        k += 1
    if k == j:
        raise ValueError(f"stoi: no conversion for '{s}'")
    return int(s2[:k])


def _parse_float(s: str) -> float:
    t = s.strip()
    low = t.lower()
    if low in ("nan", "+nan", "-nan"):
        return math.nan
    if low in ("inf", "+inf", "infinity", "+infinity"):
        return math.inf
    if low in ("-inf", "-infinity"):
        return -math.inf
    return float(t)


def read_kv_map(path: str) -> Dict[str, str]:
    """params_io.cpp:21-42: key=value lines, '#' comments, trimmed."""
    kv: Dict[str, str] = {}
    try:
        fh = open(path, "r")
    except OSError:
        raise RuntimeError(f"Cannot open params file: {path}") from None
    with fh:
        for line in fh:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            pos = line.find("=")
            if pos < 0:
                raise RuntimeError(f"Invalid key=value line in params file: {line}")
            kv[line[:pos].strip()] = line[pos + 1:].strip()
    return kv


def params_from_kv(kv: Mapping[str, str]) -> Dict[str, object]:
    """params_io.cpp:74-137: all keys required."""
    out: Dict[str, object] = {}
    for key, kind in PARAM_KEYS:
        if key not in kv:
            raise RuntimeError(f"Missing required params key: {key}")
        raw = kv[key]
        if kind == "int":
            out[key] = _parse_int(raw)
        elif kind == "float":
            out[key] = _parse_float(raw)
        elif kind == "bool":
            out[key] = _parse_bool(raw)
        else:
            out[key] = raw
    return out


def read_params_kv(path: str) -> Dict[str, object]:
    return params_from_kv(read_kv_map(path))


def _fmt(v: object) -> str:
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        if math.isnan(v):
            return "NaN"
        return "%.17g" % v
    return str(v)


def write_params_kv(path: str, params: Mapping[str, object], extra: Mapping[str, object] | None = None) -> None:
    with open(path, "w") as fh:
        for key, _ in PARAM_KEYS:
            fh.write(f"{key}={_fmt(params[key])}\n")
        for key, val in (extra or {}).items():
            fh.write(f"{key}={_fmt(val)}\n")


def to_struct(params: Mapping[str, object]) -> MpcdParams:
    """Dictionary of the 53 reference keys -> C-ABI struct."""
    s = MpcdParams()
    for key, kind in PARAM_KEYS:
        val = params[key]
        if key.startswith("boundary_"):
            setattr(s, "bc_" + key[len("boundary_"):], BC_MODES.get(str(val), -1))
        elif key == "highMode":
            s.high_mode_one_plus_coef = 1 if val == "1+coef" else 0
        elif key == "lowMode":
            s.low_mode_one_plus_coef = 1 if val == "1+coef" else 0
        elif kind == "bool":
            setattr(s, key, 1 if val else 0)
        else:
            setattr(s, key, val)
    return s


def poiseuille_params(Nx: int = 50, Ny: int = 50, a0: float = 0.2, gamma: float = 10.0, **overrides) -> Dict[str, object]:
    """Poiseuille channel parameters: periodic x, thermalising walls in y.

    Defaults are the reference's equal-benchmark set
    (build_benchmark_poiseuille_params_equal.m:30-85) with the dynamics of the
    50x50, 10-particles-per-cell channel: dt 0.05, alpha = pi/2, body force
    1e-3 and betaEOS 0.05 (betaEOS 1 is unstable at dt 0.05: the EOS kick
    scales as dt/rhoTarget, liquid_closure.cpp:428-429).
    """
    p: Dict[str, object] = {
        "Nx": Nx, "Ny": Ny, "Nc": Nx * Ny, "n": int(round(gamma * Nx * Ny)),
        "Lx": Nx * a0, "Ly": Ny * a0, "dt": 0.05, "a0": a0,
        "alpha": math.pi / 2.0, "kBT": 1.0, "g": 0.0, "bodyForceX": 1e-3,
        "useThermostat": True, "keepMeanFlow": False,
        "boundary_left": "periodic", "boundary_right": "periodic",
        "boundary_bottom": "thermalize", "boundary_top": "thermalize",
        "Utop": 0.0, "Ubottom": 0.0, "wallSigma": 1.0,
        "noSurfaceCase": True, "redistributionEnableSurfaceTopology": False,
        "redistributionWallWettingEnabled": False, "useInterfaceVelocityReorientation": False,
        "useIncompressibleRedistribution": True, "redistribAfterCollision": True, "useZoneRedistribution": True,
        "zoneTileNx": 25, "zoneTileNy": 25, "zoneUseShiftedSecondPass": True, "zonePassthroughMode": False,
        "zoneSingleFullDomainMode": False, "enableMomentumCorrectionPostRedistribution": True,
        "gamma": float(gamma), "coef": 0.1, "highMode": "coef", "lowMode": "coef", "maxRedistribPasses": 2,
        "useLocalFluidFractionThresholds": False, "fluidFracGain": 0.8, "lowThrFloor": 20.0, "highThrFloor": 25.0,
        "lowThrBulkOverride": math.nan,
        "useLiquidClosure": True, "useOptimalBetaRepair": True, "betaRepair": 0.0, "betaEOS": 0.05,
        "Kvirial": 75.0, "smoothPdriveAtInterzone": True, "smoothPdriveInterzoneWidthCells": 2,
        "smoothPdriveInterzoneBlend": 1.0, "smoothPdriveIncludeShiftedLayouts": True,
    }
    p.update(overrides)
    return p
