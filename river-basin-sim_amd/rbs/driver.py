"""Device-driven stepping: the scheduler's fast loop on the GPU.

The reference's hot loop (run.py:369-416) asks ``compute_dt`` for a step every tick,
which costs a device-to-host copy per step.  Here the host only stops the device at
sync points where it has work (outputs, slow processes); in between, librbs picks
each step with the reference's arithmetic (CFL reduction, float64 ``fill_span``
against the full sync table, rain on/off at ``rain_duration``) and replays the
steps from a captured HIP graph.
"""

from __future__ import annotations

import ctypes

import numpy as np

from rbs import _lib
from rbs.core.state import State
from rbs.scheduler import MultiRateScheduler


class DeviceRun:
    """Run parameters of one scenario, bound to its state."""

    def __init__(
        self,
        state: State,
        scheme: str,
        sched: MultiRateScheduler,
        *,
        alpha: float,
        dt_max: float,
        rain_m_s: float = 0.0,
        rain_duration: float = 0.0,
        rain_is_field: bool = False,
        rain_field_sum: float = 0.0,
        limit: bool = True,
    ):
        if scheme not in ("local_inertial", "hllc_fv"):
            raise ValueError(f"unknown scheme '{scheme}'")
        self.state = state
        self.sched = sched
        self._sync = np.ascontiguousarray(sched.sorted_sync_times(), dtype=np.float64)
        p = _lib.RunParams()
        p.scheme = _lib.SCHEME_LI if scheme == "local_inertial" else _lib.SCHEME_HLLC
        p.limit = int(bool(limit))
        p.alpha = float(alpha)
        p.dt_max = float(dt_max)
        p.rain_m_s = float(rain_m_s)
        p.rain_duration = float(rain_duration)
        p.rain_is_field = int(bool(rain_is_field))
        p.n_sync = len(self._sync)
        p.sync_times = self._sync.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        p.rain_field_sum = float(rain_field_sum)
        self.params = p

    def run_until(self, t_stop: float, max_steps: int = 0) -> int:
        """Advance the device clock to the sync time ``t_stop``; returns steps taken."""
        taken = ctypes.c_int64(0)
        _lib.call("rbs_run_until", self.state._handle, ctypes.byref(self.params),
                  float(t_stop), int(max_steps), ctypes.byref(taken))
        return int(taken.value)
