"""Local-inertial scheme: the scheme seam over librbs (reference
solver/core/local_inertial.py:515-727).

``compute_dt`` and ``step`` keep the reference's signatures, argument meaning and
error behaviour; the work is one device reduction (dt) and one fused HIP kernel
per step (``rbs_li_step``: eta, face update with Manning friction, closed edges,
donor limiter, bedload integral, continuity, Kahan sources, infiltration, open
edges), bit-for-bit the reference's float32 operation order.  A state armed with
sub-grid channels (``rbs.core.channels.arm_channels``) takes the channel forms of
the eta, face, limiter and bedload passes and of the CFL reduction (:266-500) in
the same single launch.
"""

from __future__ import annotations

import ctypes
import math

from rbs import _lib
from rbs.core.state import State


def compute_dt(state: State, alpha: float = 0.7, dt_max: float = 30.0) -> float:
    """``dt = min(alpha*dx/sqrt(g*h_max), dt_max)``; ``dt_max`` when dry (:515-544).

    With sub-grid channels ``h_max`` is the deepest water column (:487-500)."""
    out = ctypes.c_double()
    _lib.call("rbs_compute_dt", state._handle, _lib.SCHEME_LI, float(alpha), float(dt_max),
              ctypes.byref(out))
    return float(out.value)


def step(
    state: State,
    dt: float,
    rain: float = 0.0,
    limit: bool = True,
    rain_scale: float = 1.0,
    t: float = 0.0,
) -> None:
    """Advance one local-inertial step of ``dt`` seconds (:547-727)."""
    dt = float(dt)
    if not (dt > 0.0) or not math.isfinite(dt):
        raise RuntimeError(f"non-positive timestep {dt!r} at t={t}")
    _lib.call("rbs_li_step", state._handle, dt, float(rain), int(bool(limit)),
              float(rain_scale), float(t))
