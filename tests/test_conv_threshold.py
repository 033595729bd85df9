"""The IC-GN loop's convergence test without a root per iteration.

The reference continues while (float)sqrt(s2) >= conv, s2 the FP64 squared
convergence norm (src/oc_icgn.cpp:343-352).  The kernels compare s2 with an
FP64 threshold from the library (fi_conv_threshold2); these tests check that
threshold against the defining test by brute force around the boundary and
over special values, exact ties at the rounding midpoint aside."""
import ctypes

import numpy as np
import pytest

from fasticgn_amd import _native as nat


def threshold(c: float) -> float:
    out = ctypes.c_double()
    nat.call("fi_conv_threshold2", ctypes.c_float(c), ctypes.byref(out))
    return out.value


def reference_continues(s2: float, c: float) -> bool:
    with np.errstate(over="ignore"):  # a root above FLT_MAX rounds to inf, as on the device
        return bool(np.float32(np.sqrt(np.float64(s2))) >= np.float32(c))


def midpoint_below(c: float) -> float:
    """The midpoint of c and the float below it; for +inf, of FLT_MAX and
    2^128, the first root that rounds to inf."""
    c32 = np.float32(c)
    below = float(np.nextafter(c32, np.float32(0)))
    return below + 2.0 ** 103 if np.isinf(c32) else 0.5 * (below + float(c32))


def candidates(t: float, rng: np.random.Generator) -> list:
    out = [0.0, 5e-324, 1.0, np.inf, np.nan]
    if np.isfinite(t) and t > 0:
        up = down = t
        out.append(t)
        for _ in range(8):
            up = float(np.nextafter(up, np.inf))
            down = float(np.nextafter(down, 0.0))
            out += [up, down]
        for k in range(3, 16):
            out += [t * (1 + 10.0 ** -k), t * (1 - 10.0 ** -k)]
        out += list(t * rng.uniform(0.0, 4.0, 200))
    return out


CRITERIA = [1e-3, 1e-2, 0.5, 1.0, 7.0, 3e-7, 1e-30, 1e-40, 3.4e38, np.inf, 0.0, -1.0, np.nan]


@pytest.mark.parametrize("c", CRITERIA)
def test_threshold_is_the_rounded_root_test(c):
    rng = np.random.default_rng(7)
    t = threshold(c)
    m = midpoint_below(c) if c > 0 else None
    checked = 0
    for s2 in candidates(t, rng):
        if m is not None and np.isfinite(s2) and np.sqrt(np.float64(s2)) == m:
            continue  # exact tie: rounds to even, either side
        assert (s2 >= t) == reference_continues(s2, c), (c, s2, t)
        checked += 1
    assert checked >= 5


def test_threshold_special_values():
    assert threshold(0.0) == 0.0
    assert threshold(-1.0) == -np.inf
    assert np.isnan(threshold(np.nan))
    # +inf: every root from 2^128 - 2^103 on rounds to inf, so s2 from its square
    m = 2.0 ** 128 - 2.0 ** 103
    assert threshold(np.inf) == m * m
    assert reference_continues(m * m, np.inf)
    assert not reference_continues(float(np.nextafter(m, 0.0)) ** 2, np.inf)
    # the criterion itself, squared, is above the threshold; the float below is not
    c = np.float32(1e-3)
    assert float(c) ** 2 >= threshold(float(c))
    assert float(np.nextafter(c, np.float32(0))) ** 2 < threshold(float(c))
