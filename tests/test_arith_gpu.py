"""The kernels' short arithmetic routes give the long ones' bits on this device.

The friction power x^float32(7/3) (friction.py:31), the channel radius power
x^float32(4/3) (friction.py:57) and the Shields cube root x^float32(1/3)
(sediment.py:168) are evaluated correctly rounded; the Manning divisions are the IEEE
float32 quotients.  The kernels reach both through shorter routes built on the GPU's
own v_log_f32, v_exp_f32 and v_rcp_f32, guarded so that they hand every input they
cannot vouch for to the long sequences (which the oracle pins bit for bit).  Their
exactness therefore rests on that hardware, and `rbs_check_arithmetic` re-proves it on
the device running these tests:

- every float32 of [2^-11, 2^21] for each of the three powers (2.7e8 inputs each);
- v_rcp_f32's exact scaling by powers of two over 2^+-100;
- the division on every divisor significand against a sample of dividend significands
  (RBS_FULL_ARITH_CHECK=1 checks all 7.0e13 pairs, about 35 s);
- 2^31 random pairs of the ranges the callers admit, and the friction helpers against
  the plain IEEE expressions on operands drawn from the whole float32 range.
"""

from __future__ import annotations

import ctypes
import os

import pytest

pytestmark = pytest.mark.gpu


def test_short_routes_equal_the_exact_arithmetic():
    from rbs import _lib

    lib = _lib.load()
    full = os.environ.get("RBS_FULL_ARITH_CHECK") == "1"
    a_count = 1 << 23 if full else 1 << 11
    counts = (ctypes.c_uint64 * 12)()
    _lib.check(lib.rbs_check_arithmetic(a_count, counts))
    c = list(counts)
    print(f"powers: {c[10]} inputs each, left to the long route {c[3]}/{c[4]}/{c[5]}; "
          f"division pairs {c[11]}")
    assert c[0] == 0 and c[1] == 0 and c[2] == 0, f"power mismatches {c[0:3]}"
    assert c[6] == 0, f"v_rcp_f32 does not scale exactly ({c[6]} cases)"
    assert c[7] == 0, f"division significand pairs differ ({c[7]})"
    assert c[8] == 0, f"division random pairs differ ({c[8]})"
    assert c[9] == 0, f"friction helpers differ from the IEEE expressions ({c[9]})"
    # the long route is the exception: of the 32 binades of [2^-11, 2^21] the top one is
    # out of range for every power and the bottom one for 4/3 and 1/3 (the friction
    # power's depths are at least H_DRY), plus the guarded near-midpoint inputs
    n = c[10]
    for k, out in ((3, n / 32), (4, n / 16), (5, n / 16)):
        assert out <= c[k] < out + n * 1e-3
    assert c[11] == a_count << 23
