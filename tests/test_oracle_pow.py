"""The oracle's powers (x^(7/3), x^(-7/3), x^(4/3), x^(1/3), x^(3/2)) are the correctly
rounded float32 values of the exact powers.

The reference takes these powers with Warp's wp.pow, a libm powf good to about an
ulp (friction.py:57, :117, sediment.py:168, :192).  The oracle and the HIP kernels both
evaluate them by the same sequence of IEEE basic operations (so the GPU and the
oracle agree bit for bit); this pins that sequence to the mathematics: against the
exact power computed to 100 decimal digits, rounded once to float32.
"""

from __future__ import annotations

import mpmath
import numpy as np
import pytest

mpmath.mp.dps = 40


@pytest.mark.parametrize("name,expo", [("pow73", mpmath.mpf(7) / 3), ("rpow73", mpmath.mpf(-7) / 3),
                                       ("pow43", mpmath.mpf(4) / 3),
                                       ("cbrt", mpmath.mpf(1) / 3), ("pow15", mpmath.mpf(3) / 2)])
def test_powers_are_correctly_rounded(orc, name, expo):
    fn = getattr(orc, name)
    rng = np.random.default_rng(hash(name) % 2**32)
    xs = np.concatenate([(10.0 ** rng.uniform(-3.2, 4.5, 6000)).astype(np.float32),
                         np.float32([1e-3, 1.0, 2.0, 8.0, 27.0, 0.125, 1000.0, 3.0e4])])
    bad = 0
    for x in xs:
        exact = np.float32(float(mpmath.mpf(float(x)) ** expo))
        got = np.float32(fn(float(x)))
        if got != exact:
            ulps = abs(int(np.array([got]).view(np.int32)[0]) - int(np.array([exact]).view(np.int32)[0]))
            assert ulps <= 1, (x, got, exact)
            bad += 1
    assert bad == 0, f"{bad} of {len(xs)} not correctly rounded"


def _bits(x: float) -> int:
    return int(np.float32(x).view(np.uint32))


def test_every_float_of_three_binades_is_correctly_rounded(orc):
    """The seed's error repeats with period three binades (the exponent is divided by
    three), so [1, 8) holds every case the refinement meets.  Each power there is
    powl's, rounded, for all 25 165 824 floats -- but for x^(7/3) and x^(4/3) at inputs
    whose power is an exact float32 midpoint (x = s^3 with the power s^7 or s^4 of 25
    significant bits), where the float64 refinement lands within ulps of the tie and
    the side it takes is not the tie rule's.  x^(-7/3), x^(1/3) and x^(3/2) have no
    ties among float32 inputs and no exceptions at all."""
    lo, hi = _bits(1.0), _bits(8.0)
    ties = {0: [2.599609375], 3: [1.0476112365722656, 1.6093254089355469]}
    for kind in range(5):
        assert orc.pow_sweep(kind, lo, hi) == len(ties.get(kind, [])), kind
    for kind, xs in ties.items():
        num = 7 if kind == 0 else 4
        for x in xs:
            s = mpmath.cbrt(mpmath.mpf(x))  # x is a perfect cube of a dyadic rational
            assert s ** 3 == mpmath.mpf(x)
            p = s ** num  # the exact power, a dyadic rational
            f = np.float32(float(p))
            up = np.nextafter(f, np.float32(np.inf))
            dn = np.nextafter(f, np.float32(0))
            assert p in (mpmath.mpf((float(f) + float(up)) / 2), mpmath.mpf((float(f) + float(dn)) / 2))
    # and across the range the depths take, for the friction power
    assert orc.pow_sweep(1, _bits(1e-3), _bits(1e-3) + (1 << 22)) == 0
    assert orc.pow_sweep(1, _bits(900.0), _bits(900.0) + (1 << 22)) == 0


def test_manning_denominator_uses_the_seven_thirds_power(orc):
    """A single-face friction decay closes on q / (1 + g dt n^2 q / h^(7/3)) at h != 1."""
    st = orc.State(np.zeros((1, 2)), 1.0, depth=0.37, manning=0.05)
    st.qx[0, 1] = 2.0
    orc.li_step(st, 0.01, limit=False)
    expect = 2.0 / (1.0 + 9.81 * 0.01 * 0.05**2 * 2.0 / 0.37 ** (7.0 / 3.0))
    assert float(st.qx[0, 1]) == pytest.approx(expect, rel=3e-7)
