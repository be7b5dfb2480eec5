"""Morphology slow process (reference solver/processes/morphology.py:75-379).

At each activation the bed moves by the transport integral the fast loop
accumulated: ``rbs_exner_activate`` runs exner_update (float64, bounded, banked),
rebuilds ``z = f32(f64(z0) + dz_cum)`` and clears the integral while keeping its
Kahan debt -- all on the device.  The host keeps the reference's bookkeeping
(interval check, records, bounds) and, unless switched off, its per-activation
morphological Courant diagnostic, which reads a few fields back.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from rbs import _lib
from rbs.core.sediment import CourantSummary, SedimentError, courant_summary
from rbs.scheduler import EPS_T, SlowProcess


def bed_change_bounds(shape, *, alluvium_thickness=None, frozen_cells=()):
    """Per-cell bounds (lo, hi) on the cumulative bed change (morphology.py:75-110).

    By default the bed may move freely both ways.  An alluvium thickness t lets a
    cell erode at most t (t = 0: bedrock, which can still be buried), and a frozen
    cell may not change at all."""
    ny, nx = shape
    hi = np.full((ny, nx), np.inf, dtype=np.float32)
    if alluvium_thickness is None:
        lo = np.full((ny, nx), -np.inf, dtype=np.float32)
    else:
        depth = np.asarray(alluvium_thickness, dtype=np.float32)
        if depth.ndim and depth.shape != (ny, nx):
            raise SedimentError(f"the alluvium thickness field has shape {depth.shape}; the "
                                f"grid is {(ny, nx)}")
        if (depth < 0).any():
            raise SedimentError("alluvium thickness must be >= 0 m in every cell (0 is bedrock)")
        lo = np.ascontiguousarray(np.broadcast_to(-depth, (ny, nx)), dtype=np.float32)
    for i, j in frozen_cells:
        if not (0 <= i < ny and 0 <= j < nx):
            raise SedimentError(f"the frozen cell {(i, j)} lies outside the {ny}x{nx} grid")
        lo[i, j] = hi[i, j] = 0.0
    return lo, hi


@dataclass
class BedChangeRecord:
    """One activation (morphology.py:120-160): volumes of solid (m^3), the extremes of
    the cumulative bed change, and the Courant reductions when they were taken."""

    time: float
    interval_s: float
    applied_m3: float
    cumulative_m3: float
    banked_m3: float
    dz_min_m: float
    dz_max_m: float
    courant: dict = field(default_factory=dict)

    def as_dict(self) -> dict:
        out = {k: v for k, v in self.__dict__.items() if k != "courant"}
        out.update(self.courant)
        return out


class MorphologyProcess:
    """Exner on the slow clock (morphology.py:162-307)."""

    def __init__(self, state, interval_s: float, *, dz_lo=None, dz_hi=None,
                 name: str = "morphology", diagnostics: bool = True):
        sed = state.sediment
        if sed is None:
            raise SedimentError("morphology runs on a state armed by arm_sediment(); this one "
                                "is not")
        if not interval_s > 0:
            raise SedimentError(f"the morphology interval_s must be positive, got {interval_s}")
        grid_shape = state.grid.shape
        if (dz_lo is None) is not (dz_hi is None):
            raise SedimentError("bed-change bounds come as a pair: pass both dz_lo and dz_hi")
        if dz_lo is None:
            dz_lo, dz_hi = bed_change_bounds(grid_shape)
        lo = np.ascontiguousarray(dz_lo, dtype=np.float32)
        hi = np.ascontiguousarray(dz_hi, dtype=np.float32)
        if lo.shape != grid_shape or hi.shape != grid_shape:
            raise SedimentError(f"bed-change bounds of shape {lo.shape} and {hi.shape} do not "
                                f"match the grid {grid_shape}")
        if (lo > hi).any():
            raise SedimentError("the bed-change bounds cross: dz_lo exceeds dz_hi somewhere")
        self.state, self.sediment = state, sed
        self.interval_s = float(interval_s)
        self.name = name
        self.diagnostics = bool(diagnostics)
        self.records: list[BedChangeRecord] = []
        self.peak_courant = 0.0
        from rbs.core.state import DeviceField

        DeviceField(state, _lib.F_DZ_LO, grid_shape).assign(lo)
        DeviceField(state, _lib.F_DZ_HI, grid_shape).assign(hi)
        still = (lo == 0.0) & (hi == 0.0)
        self.frozen_cells = int(still.sum())
        self.floored_cells = int((np.isfinite(lo) & ~still).sum())

    def as_slow_process(self) -> SlowProcess:
        return SlowProcess(name=self.name, interval=self.interval_s, advance=self.advance)

    def advance(self, t: float, dt_slow: float) -> BedChangeRecord:
        """Apply the integral accumulated over ``dt_slow`` (one interval) to the bed."""
        if dt_slow > self.interval_s * (1.0 + 1e-6) + EPS_T:
            raise SedimentError(
                f"{dt_slow} s handed to a morphology process whose interval is "
                f"{self.interval_s} s: several intervals would have collapsed into one bed "
                "update")
        st, sed, g = self.state, self.sediment, self.state.grid
        before = sed.totals()["sum"]
        dz_before = sed.dz_cum.numpy() if self.diagnostics else None
        _lib.call("rbs_exner_activate", st._handle, float(sed.porosity))
        per_cell = g.cell_area * (1.0 - sed.porosity)
        tot = sed.totals()  # reduced on the device: the bed-change fields stay there
        courant: dict = {}
        if self.diagnostics:
            cs: CourantSummary = courant_summary(st, sed.dz_cum.numpy() - dz_before,
                                                 float(dt_slow))
            self.peak_courant = max(self.peak_courant, cs.courant)
            courant = cs.as_dict()
        rec = BedChangeRecord(time=float(t), interval_s=float(dt_slow),
                              applied_m3=(tot["sum"] - before) * per_cell,
                              cumulative_m3=tot["sum"] * per_cell,
                              banked_m3=tot["unapplied"] * per_cell,
                              dz_min_m=tot["min"], dz_max_m=tot["max"], courant=courant)
        self.records.append(rec)
        return rec

    @property
    def series(self) -> list[dict]:
        return [r.as_dict() for r in self.records]

    def as_attrs(self) -> dict:
        return {**self.sediment.as_attrs(), "interval_s": self.interval_s,
                "frozen_cells": self.frozen_cells, "floored_cells": self.floored_cells,
                "peak_courant": self.peak_courant}
