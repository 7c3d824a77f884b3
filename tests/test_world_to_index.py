"""The kernels' test for "fmod(c, res) == 0" in the rounded world->index
(VDBMapping.hpp:612-631; vdb_mapping_amd/csrc/upd_common.h world_to_index1):
fma(-k, res, c) == 0 with k = rint(c * (1 / res)). fma's result is zero exactly
when c - k res is zero in exact arithmetic, so the test is restated here with
exact rationals and compared with math.fmod over exact multiples, near misses
and random coordinates."""
import math
import random
from fractions import Fraction


def _kernel_is_multiple(c, res):
    inv = 1.0 / res
    k = round(c * inv)  # rint: round half to even, as Python's round
    return Fraction(c) - k * Fraction(res) == 0


def _cases(res, rng):
    out = [0.0, -0.0, res, -res, 1e6 * res, -7.0 * res]
    for _ in range(3000):
        k = rng.randint(-2_000_000, 2_000_000)
        c = k * res  # rounded product: sometimes an exact multiple, sometimes not
        out += [c, math.nextafter(c, math.inf), math.nextafter(c, -math.inf), c + res / 2]
        out.append(rng.uniform(-100.0, 100.0))
    return out


def test_multiple_test_equals_fmod():
    rng = random.Random(7)
    checked = multiples = 0
    for res in (0.05, 0.1, 0.02, 0.2, 0.25, 0.125, 0.5, 1.0, 0.3, 1.0 / 3.0, 0.07):
        for c in _cases(res, rng):
            ref = math.fmod(c, res) == 0.0
            assert _kernel_is_multiple(c, res) == ref, (c, res)
            checked += 1
            multiples += ref
    assert checked > 100_000 and multiples > 1000  # both outcomes exercised
