"""Host logic check: the GPU library's kernel bodies (mpcd_core.h, host build)
against the oracle's counter mode on the golden cases -- bitwise."""
import numpy as np
import pytest

from src_incompressible_openmp_amd.params import LAYOUT_BASE, LAYOUT_SHIFTED, to_struct
from tests.golden.cases import SMALL_CASES, case_params, case_state
from tests.oracle_lib import CTR


@pytest.fixture(scope="module")
def core():
    from tests.core_host_lib import CoreHost

    return CoreHost()


@pytest.fixture(scope="module")
def core_small():
    from tests.core_host_lib import CoreHost

    return CoreHost(small=True)


def _bits(a, b):
    return np.ascontiguousarray(a).tobytes() == np.ascontiguousarray(b).tobytes()


@pytest.mark.parametrize("case", SMALL_CASES)
def test_bc_matches_oracle(core, oracle, case):
    p = to_struct(case_params(case))
    x, v, _, _ = case_state(case)
    xo, vo = x.copy(), v.copy()
    oracle.force_stream(p, xo, vo)
    oracle.apply_bc(CTR, p, xo, vo, 104729)
    xc, vc = x.copy(), v.copy()
    assert core.force_stream_bc(p, xc, vc, 104729) == 0
    assert _bits(xc, xo) and _bits(vc, vo)


@pytest.mark.parametrize("case", SMALL_CASES)
def test_collide_matches_oracle(core, oracle, case):
    p = to_struct(case_params(case))
    x, v, _, _ = case_state(case)
    cid, _ = oracle.shift_cell_ids(CTR, p, x, 130363)
    vo, vc = v.copy(), v.copy()
    oracle.collide(CTR, p, vo, cid, 433494437)
    core.collide(p, vc, cid, 433494437)
    assert _bits(vc, vo)


# working-set variants of the zone kernel: 16-bit indices with the links in
# the working set, 16-bit with every link in the global region, 32-bit
VARIANTS = {"u16": (False, 38 * 1024), "u16_spill": (False, 0), "i32": (True, 64 * 1024)}


@pytest.mark.parametrize("variant", sorted(VARIANTS))
@pytest.mark.parametrize("case", SMALL_CASES + ("cfg0",))
@pytest.mark.parametrize("layout", [LAYOUT_BASE, LAYOUT_SHIFTED])
def test_zone_pass_matches_oracle(core, oracle, case, layout, variant):
    p = to_struct(case_params(case))
    x, v, _, _ = case_state(case)
    # run on a post-collision state (the pass's real input)
    oracle.force_stream(p, x, v)
    oracle.apply_bc(CTR, p, x, v, 104729)
    cid, _ = oracle.shift_cell_ids(CTR, p, x, 130363)
    oracle.collide(CTR, p, v, cid, 433494437)
    xo, vo = x.copy(), v.copy()
    m, _ = oracle.zone_pass(CTR, p, xo, vo, layout, 88172645463325252)
    xc, vc = x.copy(), v.copy()
    wide, ws = VARIANTS[variant]
    moved, corr = core.zone_pass(p, xc, vc, layout, 88172645463325252, wide=wide, ws_bytes=ws)
    assert moved[0] == m["nParticlesMovedDense"] and moved[1] == m["nParticlesMovedSparse"]
    assert m["nParticlesMovedDense"] + m["nParticlesMovedSparse"] > 0
    assert _bits(xc, xo)
    assert _bits(vc, vo)
    assert corr[0] == m["corrRmsDU"] and corr[1] == m["corrMaxDU"]


@pytest.mark.parametrize("case", SMALL_CASES + ("cfg0",))
@pytest.mark.parametrize("layout", [LAYOUT_BASE, LAYOUT_SHIFTED])
def test_zone_pass_past_staging_limits(core_small, oracle, case, layout):
    """2 staged dense movers and 4-entry LDS donor lists: the unstaged mover
    path and the global donor-list region give the same result."""
    p = to_struct(case_params(case))
    x, v, _, _ = case_state(case)
    oracle.force_stream(p, x, v)
    oracle.apply_bc(CTR, p, x, v, 104729)
    cid, _ = oracle.shift_cell_ids(CTR, p, x, 130363)
    oracle.collide(CTR, p, v, cid, 433494437)
    xo, vo = x.copy(), v.copy()
    m, _ = oracle.zone_pass(CTR, p, xo, vo, layout, 88172645463325252)
    xc, vc = x.copy(), v.copy()
    moved, corr = core_small.zone_pass(p, xc, vc, layout, 88172645463325252)
    assert moved[0] == m["nParticlesMovedDense"] and moved[1] == m["nParticlesMovedSparse"]
    assert _bits(xc, xo) and _bits(vc, vo)
    assert corr[0] == m["corrRmsDU"] and corr[1] == m["corrMaxDU"]
