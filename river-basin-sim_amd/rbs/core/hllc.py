"""Well-balanced HLLC finite-volume scheme: the scheme seam over librbs (reference
solver/core/hllc.py:972-1373).

``compute_dt`` reduces the wave speed on the device; ``step`` runs the SSP-RK2
step (MUSCL-minmod + Audusse hydrostatic reconstruction + HLLC flux, per-edge
ghost boundaries, donor mass limiter, friction, sources) as fused HIP kernels.
"""

from __future__ import annotations

import ctypes
import math

from rbs import _lib
from rbs.core.state import State


def arm_hllc(state: State) -> None:
    """Allocate momentum and stage buffers (idempotent; hllc.py:123-152)."""
    state._arm(_lib.ARM_HLLC)


def compute_dt(state: State, alpha: float = 0.45, dt_max: float = 30.0) -> float:
    """``dt = min(alpha*dx/max(|u|,|v|+sqrt(g h)), dt_max)`` (hllc.py:972-990)."""
    arm_hllc(state)
    out = ctypes.c_double()
    _lib.call("rbs_compute_dt", state._handle, _lib.SCHEME_HLLC, float(alpha), float(dt_max),
              ctypes.byref(out))
    return float(out.value)


def step(
    state: State,
    dt: float,
    rain: float = 0.0,
    rain_scale: float = 1.0,
    t: float = 0.0,
) -> None:
    """Advance one HLLC step (hllc.py:1272-1373); stages read at ``t + dt/2``."""
    arm_hllc(state)
    dt = float(dt)
    if not (dt > 0.0) or not math.isfinite(dt):
        raise RuntimeError(f"non-positive timestep {dt!r} at t={t}")
    _lib.call("rbs_hllc_step", state._handle, dt, float(rain), float(rain_scale), float(t))
