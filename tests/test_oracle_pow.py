"""The oracle's powers are the correctly rounded float32 values of x^e with the
reference's float32 exponents, and its friction term divides by the power.

The reference takes its powers with Warp's wp.pow on float32 arguments, with the
exponents written as Python floats: `wp.pow(h_flow, 7.0 / 3.0)` (friction.py:31),
`wp.pow(radius, 4.0 / 3.0)` (friction.py:57), `wp.pow(radius, 1.0 / 3.0)`
(sediment.py:168) and `wp.pow(excess, 1.5)` (sediment.py:192).  The exponent reaches
the float32 pow as float32(7/3) = 2.3333332538604736, float32(4/3), float32(1/3) and
1.5.  A libm powf is good to about an ulp; the oracle's default mode (and the HIP
kernels, by the same operation sequence) gives the correctly rounded power.  These
tests pin that mode to the mathematics (40-digit mpmath powers, rounded once) and to
powl over every float32 of the ranges the solver's depths take, and pin the libm
mode to glibc's powf with the same exponents.
"""

from __future__ import annotations

import mpmath
import numpy as np
import pytest

mpmath.mp.dps = 50

E73 = float(np.float32(7.0 / 3.0))
E43 = float(np.float32(4.0 / 3.0))
E13 = float(np.float32(1.0 / 3.0))


def _bits(x: float) -> int:
    return int(np.float32(x).view(np.uint32))


def _rn32(v) -> np.float32:
    """Round an mpmath value to the nearest float32 (ties to even), subnormals included."""
    if v == 0:
        return np.float32(0.0)
    f = np.float32(float(v))
    cands = [np.nextafter(f, np.float32(0)), f, np.nextafter(f, np.float32(np.inf))]
    return min(cands, key=lambda c: (abs(mpmath.mpf(float(c)) - v),
                                     int(np.array([c]).view(np.uint32)[0]) & 1))


def test_the_exponents_are_the_float32_ones():
    """What a Python float exponent becomes in a float32 pow."""
    assert E73 == 2.3333332538604736 and E73 < 7.0 / 3.0
    assert E43 == 1.3333333730697632 and E43 > 4.0 / 3.0
    assert E13 == 0.3333333432674408 and E13 > 1.0 / 3.0


@pytest.mark.parametrize("name,expo", [("pow73", E73), ("pow43", E43), ("cbrt", E13), ("pow15", 1.5)])
def test_powers_are_correctly_rounded(orc, name, expo):
    """Random inputs over the whole float32 range (subnormal inputs and results
    included) against the 50-digit power rounded once."""
    fn = getattr(orc, name)
    rng = np.random.default_rng(sum(map(ord, name)))
    xs = np.concatenate([(10.0 ** rng.uniform(-3.2, 4.5, 4000)).astype(np.float32),
                         (10.0 ** rng.uniform(-44.5, 38.0, 2000)).astype(np.float32),
                         np.float32([1e-3, 1.0, 2.0, 8.0, 27.0, 0.125, 1000.0, 3.0e4, 0.37])])
    e = mpmath.mpf(expo)
    checked = 0
    for x in xs[xs > 0]:
        exact = mpmath.mpf(float(x)) ** e
        if exact > 3.4e38:
            continue
        got = np.float32(fn(float(x)))
        assert got == _rn32(exact), (float(x), float(got), float(_rn32(exact)))
        checked += 1
    assert checked > 5000


def test_the_power_is_not_the_exact_rational_one(orc):
    """A plausible mistake -- the exponent 7/3 itself -- moves most results: at h = 0.37
    x^(7/3) and x^float32(7/3) are different float32 numbers, and the oracle has the
    latter."""
    h = 0.37
    want = _rn32(mpmath.mpf(float(np.float32(h))) ** mpmath.mpf(E73))
    wrong = _rn32(mpmath.mpf(float(np.float32(h))) ** (mpmath.mpf(7) / 3))
    assert want != wrong
    assert np.float32(orc.pow73(h)) == want


def test_ln_inside_the_powers_is_good_to_3e8(orc):
    rng = np.random.default_rng(11)
    xs = np.concatenate([(10.0 ** rng.uniform(-37.5, 38.0, 3000)).astype(np.float32),
                         np.float32([1.0, 2.0, 0.5, 1.41421354, 1.41421366, 0.70710677])])
    worst = max(abs(orc.ln_coarse(float(x)) - float(mpmath.log(mpmath.mpf(float(x))))) for x in xs)
    assert worst < 3e-8


# Misses against powl over [1e-3, 1e4): every one is a case where the exact power lies
# within 1e-15 of a rounding midpoint (the float64 evaluation takes the other side), or
# where powl itself misrounds (long double, then float: a double rounding).
_KNOWN_MISSES = {0: [199.05653381347656], 1: [], 2: [1.9799857139587402]}


@pytest.mark.parametrize("kind,p", [(0, 7), (1, 4), (2, 1)])
def test_every_float_of_the_depth_range_against_powl(orc, kind, p):
    """All 194.6 M float32 in [1e-3, 1e4) against powl with the same float32 exponent,
    rounded; the handful of differences are each resolved by the 50-digit power."""
    lo, hi = _bits(1e-3), _bits(1e4)
    known = _KNOWN_MISSES[kind]
    assert orc.pow_sweep(kind, lo, hi) == len(known)
    fn = (orc.pow73, orc.pow43, orc.cbrt)[kind]
    e = mpmath.mpf(float(np.float32(p / 3.0)))
    for x in known:
        assert orc.pow_sweep(kind, _bits(x), _bits(x) + 1) == 1
        exact = mpmath.mpf(x) ** e
        got = np.float32(fn(x))
        if got == _rn32(exact):
            continue  # powl misrounded
        other = np.nextafter(got, np.float32(np.inf) if exact > mpmath.mpf(float(got)) else np.float32(0))
        mid = (mpmath.mpf(float(got)) + mpmath.mpf(float(other))) / 2
        assert abs(exact - mid) / exact < 1e-15, x


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_the_libm_mode_is_glibc_powf_with_float32_exponents(orc, kind):
    """glibc's powf with the float32 exponent differs from the correctly rounded power
    on well under 0.1 % of the depth range; with the rational exponent it would differ
    on most of it.  This pins what the libm mode evaluates."""
    lo, hi = _bits(1.0), _bits(8.0)
    bad = orc.pow_sweep(kind, lo, hi, libm=True)
    assert bad < 1e-3 * (hi - lo)
    with orc.pow_mode("libm"):
        x = 0.37
        assert np.float32((orc.pow73, orc.pow43, orc.cbrt)[kind](x)) == np.float32(
            np.float32(x) ** np.float32((E73, E43, E13)[kind]))  # numpy's float32 power is libm's


def _f32_manning_decay(q, h, n, dt, g=9.81, power=None, reciprocal=False):
    """friction.py:31 in float32, left to right, with the given power of h."""
    f = np.float32
    pw = power(h)
    t = f(f(f(f(f(g) * f(dt)) * f(n)) * f(n)) * f(abs(q)))
    frac = t * (f(1.0) / pw) if reciprocal else t / pw
    return f(f(q) / f(f(1.0) + frac))


def test_manning_denominator_divides_by_the_float32_power(orc):
    """A single frictional face on flat water closes bitwise on q / (1 + g dt n^2 |q| /
    h^float32(7/3)) in float32; the exact-rational power or a multiplication by the
    reciprocal gives different bits for some of these faces."""
    cr = lambda e: (lambda h: _rn32(mpmath.mpf(float(np.float32(h))) ** e))
    differs_power = differs_recip = 0
    for h, q, n, dt in [(0.37, 2.0, 0.05, 0.01), (0.0123, 0.004, 0.08, 1.5), (3.7, 9.0, 0.03, 0.7),
                        (0.0021, 0.0007, 0.1, 2.0), (12.5, 40.0, 0.04, 0.9), (0.051, 0.03, 0.06, 1.1)]:
        st = orc.State(np.zeros((1, 2)), 1.0, depth=h, manning=n)
        st.qx[0, 1] = q
        orc.li_step(st, dt, limit=False)
        want = _f32_manning_decay(q, h, n, dt, power=cr(mpmath.mpf(E73)))
        assert np.float32(st.qx[0, 1]) == want, (h, q)
        differs_power += _f32_manning_decay(q, h, n, dt, power=cr(mpmath.mpf(7) / 3)) != want
        differs_recip += _f32_manning_decay(q, h, n, dt, power=cr(mpmath.mpf(E73)), reciprocal=True) != want
    assert differs_power >= 1 and differs_recip >= 1
