"""Zones that are not rectangles on the GPU (run_zone_kernel_inplace with a
ZoneDescriptor.maskVec of any shape, zone_pass.cpp:1065-1081): a rectangle
with a hole, an L, scattered cells and a band across the periodic x
boundary, on three grids, against the oracle's masked zone kernel with the
same counter draws.  State, stats and moved counts bit for bit; diag within
the tolerances of the rectangular zones' diagnostics test.

The oracle's masked zones are pinned to the compiled original through the
same masks (tests/test_zone_kernel_golden.py), and a mask that fills its
bounding rectangle is checked against the rectangle entry point."""
import os

import numpy as np
import pytest

from src_incompressible_openmp_amd.params import LAYOUT_BASE, to_struct
from tests.golden.cases import case_params, case_state
from tests.oracle_lib import CTR
from tests.test_gpu_diag import diag_close

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "zonekernel.npz")
CASES = ("small", "override", "cfg0")
MASKS = ("notch", "ell", "scatter", "wrap")


@pytest.fixture(scope="module")
def gpu():
    from src_incompressible_openmp_amd import gpu as g

    g.load()
    assert g.device_count() >= 1, "no GPU visible"
    return g


@pytest.fixture(scope="module")
def masks():
    return np.load(FIX)


@pytest.mark.parametrize("case", CASES)
def test_masked_zones_match_oracle(gpu, oracle, masks, case):
    p = to_struct(case_params(case))
    x, v, _, _ = case_state(case)
    scale_p = float(np.abs(v).sum())
    scale_e = float(0.5 * np.sum(v * v))
    for mi, name in enumerate(MASKS):
        m = masks[f"{case}_mask_{name}"]
        seed = 5101 + mi
        print(f"masked zone: {case} {name} ({int(m.sum())} cells)", flush=True)
        xo, vo = x.copy(), v.copy()
        st, dgo, mv = oracle.zone_kernel_mask(CTR, p, xo, vo, m, seed)
        xg, vg = x.copy(), v.copy()
        sg, mg, dgg = gpu.run_zone_kernel_mask(p, xg, vg, m, seed, diag=True)
        assert mv.sum() > 0, name
        assert xg.tobytes() == xo.tobytes() and vg.tobytes() == vo.tobytes(), name
        assert np.array_equal(mg, mv) and sg.tobytes() == st.tobytes(), name
        diag_close(dgg, dgo, scale_p, scale_e)


def test_full_mask_is_the_rectangle(gpu, masks):
    """A mask that fills a layout zone's rectangle gives the rectangle's
    results."""
    p = to_struct(case_params("override"))
    x, v, _, _ = case_state("override")
    _, rects = gpu.build_zone_layout(p, LAYOUT_BASE)
    xs, xe, ys, ye = (int(t) for t in rects[1])
    m = np.zeros((p.Nx, p.Ny), np.uint8)
    m[xs:xe, ys:ye] = 1
    xa, va = x.copy(), v.copy()
    sa, ma = gpu.run_zone_kernel(p, xa, va, rects[1], 77)
    xb, vb = x.copy(), v.copy()
    sb, mb = gpu.run_zone_kernel_mask(p, xb, vb, m.reshape(-1), 77)
    assert xa.tobytes() == xb.tobytes() and va.tobytes() == vb.tobytes()
    assert sa.tobytes() == sb.tobytes() and np.array_equal(ma, mb)


def test_empty_mask_is_refused(gpu):
    p = to_struct(case_params("small"))
    x, v, _, _ = case_state("small")
    with pytest.raises(gpu.MpcdError):
        gpu.run_zone_kernel_mask(p, x.copy(), v.copy(), np.zeros(p.Nc, np.uint8), 1)
