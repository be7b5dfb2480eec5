"""Inflow hydrographs -- prescribed discharge point sources (reference
solver/processes/inflow.py:116-200).

The injector registers its sources with the state's device table
(``rbs_set_inflows``).  ``apply`` keeps the reference's host-driven contract
(midpoint discharge, compensated float32 add, returns the float32-requested
volume for the ledger); a device-driven run evaluates the same hydrographs on the
device every step and banks the volume in the device ledger itself.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from rbs import _lib


@dataclass
class Inflow:
    """A point-source hydrograph (config.py:95-121)."""

    cell: tuple[int, int]
    hydrograph: list[tuple[float, float]]

    @property
    def breakpoints(self) -> list[float]:
        return [t for t, _ in self.hydrograph]

    def discharge_at(self, t: float) -> float:
        hg = self.hydrograph
        if not hg or t < hg[0][0] or t > hg[-1][0]:
            return 0.0
        for (t0, q0), (t1, q1) in zip(hg, hg[1:]):
            if t0 <= t <= t1:
                if t1 == t0:
                    return q1
                return q0 + (q1 - q0) * (t - t0) / (t1 - t0)
        return hg[-1][1]


class InflowInjector:
    """Applies a set of hydrographs each step (inflow.py:116-200)."""

    def __init__(self, inflows: list[Inflow], state):
        self.inflows = list(inflows)
        self.state = state
        self.cell_area = state.grid.cell_area
        rows = np.ascontiguousarray([i.cell[0] for i in self.inflows], dtype=np.int32)
        cols = np.ascontiguousarray([i.cell[1] for i in self.inflows], dtype=np.int32)
        offs = np.zeros(len(self.inflows) + 1, np.int32)
        kt, kq = [], []
        for k, inf in enumerate(self.inflows):
            kt.extend(float(t) for t, _ in inf.hydrograph)
            kq.extend(float(q) for _, q in inf.hydrograph)
            offs[k + 1] = len(kt)
        kt = np.ascontiguousarray(kt if kt else [0.0], dtype=np.float64)
        kq = np.ascontiguousarray(kq if kq else [0.0], dtype=np.float64)
        ip = ctypes.POINTER(ctypes.c_int32)
        dp = ctypes.POINTER(ctypes.c_double)
        _lib.call("rbs_set_inflows", state._handle, len(self.inflows), rows.ctypes.data_as(ip),
                  cols.ctypes.data_as(ip), offs.ctypes.data_as(ip), kt.ctypes.data_as(dp),
                  kq.ctypes.data_as(dp))

    def breakpoints(self) -> list[float]:
        return sorted({t for inf in self.inflows for t in inf.breakpoints})

    def apply(self, state, t: float, dt: float) -> float:
        """Inject the midpoint discharge over [t, t+dt]; return the requested volume (m^3)."""
        if not self.inflows:
            return 0.0
        out = ctypes.c_double()
        _lib.call("rbs_inflow_apply", state._handle, float(t), float(dt), ctypes.byref(out))
        return float(out.value)
