"""Pin the counter-based generator the oracle's counter mode and the GPU
library share (include/mpcd_rng.h): published Philox4x32-10 known answers,
exact bit constructions, and the portable log / sincos / Box-Muller against
libm on a sweep."""
import math

import numpy as np
import pytest


# Random123 known-answer vectors for philox4x32 with 10 rounds:
# (counter words, key words, expected output words)
KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF),
     (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,want", KAT)
def test_philox_known_answers(oracle, ctr, key, want):
    seed = key[0] | (key[1] << 32)  # key = (low, high 32 bits of the seed)
    got = oracle.ctr_philox(seed, ctr)
    assert [int(v) for v in got] == list(want)


def test_log_matches_libm(oracle):
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.uniform(1e-300, 1.0, 2000), rng.uniform(0.5, 2.0, 2000),
                         np.exp(rng.uniform(-700, 700, 2000)), [1.0, 2.0, 0.5, math.sqrt(2.0), 2.0 ** -1022, 2.0 ** -53]])
    worst = 0.0
    for x in xs:
        ref = math.log(x)
        got = oracle.ctr_log(float(x))
        ulp = math.ulp(ref) if ref != 0.0 else math.ulp(1.0)
        worst = max(worst, abs(got - ref) / ulp)
    assert oracle.ctr_log(1.0) == 0.0
    assert worst <= 2.0, worst


def test_sincos_2pi_matches_libm(oracle):
    rng = np.random.default_rng(8)
    us = np.concatenate([rng.uniform(0.0, 1.0, 3000), [0.0, 0.25, 0.5, 0.75, 1.0 - 2.0 ** -53]])
    # libm's argument 2*pi*u is itself rounded (up to half an ulp of ~6.3),
    # so the comparison allows a few units of 1e-16
    for u in us:
        s, c = oracle.ctr_sincos_2pi(float(u))
        assert abs(s - math.sin(2.0 * math.pi * u)) <= 2e-15
        assert abs(c - math.cos(2.0 * math.pi * u)) <= 2e-15
        assert abs(s * s + c * c - 1.0) <= 4.5e-16


def u01(hi, lo):
    return float(((int(hi) << 32) | int(lo)) >> 11) * 2.0 ** -53


def test_normal_pair_is_box_muller_of_the_stream(oracle):
    """The wall draw of (pid, sweep, wall) is Box-Muller of the first two
    53-bit uniforms of its Philox block (counter tag 0x01)."""
    seed = 104729 + 10007 * 3
    for pid, sweep, wall in [(0, 0, 0), (17, 1, 2), (123456, 3, 3), (9, 2, 1)]:
        w = oracle.ctr_philox(seed, ((0x01 << 24) | (sweep * 4 + wall), pid, 0, 0))
        a, b = u01(w[0], w[1]), u01(w[2], w[3])
        n1, n2 = oracle.ctr_normal_pair(seed, pid, sweep, wall)
        r = math.sqrt(-2.0 * math.log(1.0 - a))
        assert abs(n1 - r * math.cos(2.0 * math.pi * b)) <= 1e-14 * max(1.0, abs(n1))
        assert abs(n2 - r * math.sin(2.0 * math.pi * b)) <= 1e-14 * max(1.0, abs(n2))


def test_normal_pair_moments(oracle):
    vals = []
    for pid in range(4000):
        vals.extend(oracle.ctr_normal_pair(99, pid, 0, 2))
    v = np.array(vals)
    assert abs(v.mean()) < 0.05 and abs(v.std() - 1.0) < 0.05
    assert abs(np.mean(v ** 4) - 3.0) < 0.3
