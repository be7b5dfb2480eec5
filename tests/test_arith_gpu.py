"""The local-inertial face's short arithmetic routes give the long ones' bits here.

The LI face (local_inertial.py:72-129 with friction.py:18-31) evaluates the friction
power h^float32(7/3) correctly rounded and its two divisions as the IEEE float32
quotients.  The kernel reaches both through shorter routes built on the GPU's own
v_log_f32, v_exp_f32 and v_rcp_f32, guarded so that they hand every operand they
cannot vouch for to the long sequences (which the oracle pins bit for bit).  Their
exactness therefore rests on that hardware, and `rbs_check_arithmetic` re-proves it on
the device running these tests:

- every float32 of [2^-11, 2^21] for the power (2.7e8 inputs);
- v_rcp_f32's exact scaling by powers of two over 2^+-100;
- the division on every divisor significand against a sample of dividend significands
  (RBS_FULL_ARITH_CHECK=1 checks all 7.0e13 pairs, about 35 s);
- 2^31 random pairs of the face's division ranges, and the whole face update against
  the plain IEEE expression on operands drawn from the whole float32 range.
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
    counts = (ctypes.c_uint64 * 8)()
    _lib.check(lib.rbs_check_arithmetic(a_count, counts))
    c = list(counts)
    print(f"power: {c[6]} inputs, left to the long route {c[1]}; division pairs {c[7]}")
    assert c[0] == 0, f"power mismatches ({c[0]})"
    assert c[2] == 0, f"v_rcp_f32 does not scale exactly ({c[2]} cases)"
    assert c[3] == 0, f"division significand pairs differ ({c[3]})"
    assert c[4] == 0, f"division random pairs differ ({c[4]})"
    assert c[5] == 0, f"face updates differ from the IEEE expression ({c[5]})"
    # the long route is the exception: the top binade of [2^-11, 2^21] is out of range
    # (1/32 of the inputs), plus the guarded near-midpoint inputs
    n = c[6]
    assert n / 32 <= c[1] < n / 32 + n * 1e-3
    assert c[7] == a_count << 23
