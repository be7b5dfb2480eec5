"""Morphology slow process (reference solver/processes/morphology.py:75-379).

At each activation the bed moves by the transport integral the fast loop
accumulated: ``rbs_exner_activate`` runs exner_update (float64, bounded, banked),
rebuilds ``z = f32(f64(z0) + dz_cum)`` and clears the integral while keeping its
Kahan debt -- all on the device.  The host keeps the reference's bookkeeping
(interval check, records, bounds).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from rbs import _lib
from rbs.core.sediment import SedimentError
from rbs.scheduler import EPS_T, SlowProcess


def bed_change_bounds(shape, *, alluvium_thickness=None, frozen_cells=()):
    """morphology.py:75-110: per-cell bounds on the cumulative bed change."""
    ny, nx = shape
    lo = np.full((ny, nx), -np.inf, dtype=np.float32)
    hi = np.full((ny, nx), np.inf, dtype=np.float32)
    if alluvium_thickness is not None:
        thickness = np.asarray(alluvium_thickness, dtype=np.float32)
        if thickness.ndim and thickness.shape != (ny, nx):
            raise SedimentError(f"alluvium thickness shape {thickness.shape} != grid {(ny, nx)}")
        if (thickness < 0).any():
            raise SedimentError("alluvium thickness must be >= 0 m (0 = bedrock, not a hole)")
        lo = np.broadcast_to(-thickness, (ny, nx)).astype(np.float32).copy()
    for i, j in frozen_cells:
        if not (0 <= i < ny and 0 <= j < nx):
            raise SedimentError(f"frozen cell {(i, j)} is outside the {ny}x{nx} grid")
        lo[i, j] = 0.0
        hi[i, j] = 0.0
    return lo, hi


@dataclass
class BedChangeRecord:
    time: float
    interval_s: float
    applied_m3: float
    cumulative_m3: float
    banked_m3: float
    dz_min_m: float
    dz_max_m: float

    def as_dict(self) -> dict:
        return dict(self.__dict__)


class MorphologyProcess:
    """Applies Exner on the slow clock (morphology.py:162-307)."""

    def __init__(self, state, interval_s: float, *, dz_lo=None, dz_hi=None,
                 name: str = "morphology"):
        sed = state.sediment
        if sed is None:
            raise SedimentError("morphology needs a state armed with arm_sediment()")
        if interval_s <= 0:
            raise SedimentError(f"morphology interval_s must be > 0, got {interval_s}")
        self.state, self.sediment = state, sed
        self.interval_s = float(interval_s)
        self.name = name
        self.records: list[BedChangeRecord] = []
        shape = state.grid.shape
        if (dz_lo is None) != (dz_hi is None):
            raise SedimentError("give both dz_lo and dz_hi, or neither")
        if dz_lo is None:
            dz_lo, dz_hi = bed_change_bounds(shape)
        lo = np.ascontiguousarray(dz_lo, dtype=np.float32)
        hi = np.ascontiguousarray(dz_hi, dtype=np.float32)
        if lo.shape != shape or hi.shape != shape:
            raise SedimentError(f"bed-change bounds {lo.shape}/{hi.shape} != grid {shape}")
        if (lo > hi).any():
            raise SedimentError("bed-change bounds cross: dz_lo must be <= dz_hi everywhere")
        from rbs.core.state import DeviceField

        DeviceField(state, _lib.F_DZ_LO, shape).assign(lo)
        DeviceField(state, _lib.F_DZ_HI, shape).assign(hi)
        frozen = (lo == 0.0) & (hi == 0.0)
        self.frozen_cells = int(frozen.sum())
        self.floored_cells = int((np.isfinite(lo) & ~frozen).sum())

    def as_slow_process(self) -> SlowProcess:
        return SlowProcess(name=self.name, interval=self.interval_s, advance=self.advance)

    def advance(self, t: float, dt_slow: float) -> BedChangeRecord:
        if dt_slow > self.interval_s * (1.0 + 1e-6) + EPS_T:
            raise SedimentError(
                f"morphology was handed {dt_slow} s to apply but is configured for "
                f"{self.interval_s} s: more than one interval has collapsed into a single "
                "bed update")
        st, sed, g = self.state, self.sediment, self.state.grid
        before = sed.totals()["sum"]
        _lib.call("rbs_exner_activate", st._handle, float(sed.porosity))
        solid = g.cell_area * (1.0 - sed.porosity)
        tot = sed.totals()  # device reductions: the bed-change fields stay on the GPU
        cumulative = tot["sum"]
        rec = BedChangeRecord(time=float(t), interval_s=float(dt_slow),
                              applied_m3=(cumulative - before) * solid,
                              cumulative_m3=cumulative * solid,
                              banked_m3=tot["unapplied"] * solid,
                              dz_min_m=tot["min"], dz_max_m=tot["max"])
        self.records.append(rec)
        return rec

    @property
    def series(self) -> list[dict]:
        return [r.as_dict() for r in self.records]

    def as_attrs(self) -> dict:
        return {**self.sediment.as_attrs(), "interval_s": self.interval_s,
                "frozen_cells": self.frozen_cells, "floored_cells": self.floored_cells}
