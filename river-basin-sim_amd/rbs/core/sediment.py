"""Bedload transport (Meyer-Peter-Mueller) and the Exner bed update
(reference solver/core/sediment.py).

The transport integral is accumulated inside the fused local-inertial kernel
(after the limiter, before continuity, exactly where the reference places it,
local_inertial.py:680-687); the Exner update and bed rebuild run at morphology
activations (``rbs_exner_activate``).  This module arms the state, validates
inputs and keeps the reference's float64 host references for the law.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from rbs import _lib
from rbs.core.grid import GRAVITY, H_DRY

SUBMERGED_SG = 1.65  # sediment.py:146
SHIELDS_CRITICAL = 0.047  # :148
MPM_COEFFICIENT = 8.0  # :150
MPM_EXPONENT = 1.5  # :151
MORPH_COURANT_GATE = 1.0


class SedimentError(ValueError):
    """Sediment inputs are inconsistent with the grid or with themselves."""


class SedimentState:
    """Device-side morphology fields of one state (sediment.py:548-616)."""

    def __init__(self, state, porosity: float, d50_min: float, d50_max: float, zero_cells: int):
        self.state = state
        self.porosity = float(porosity)
        self.d50_min, self.d50_max, self.zero_d50_cells = d50_min, d50_max, zero_cells

    def _f(self, fid, shape=None, dtype=np.float32):
        from rbs.core.state import DeviceField

        return DeviceField(self.state, fid, shape or self.state.grid.shape, dtype)

    @property
    def d50(self):
        return self._f(_lib.F_D50)

    @property
    def z0(self):
        return self._f(_lib.F_Z0)

    @property
    def dz_cum(self):
        return self._f(_lib.F_DZ_CUM, dtype=np.float64)

    @property
    def dz_unapplied(self):
        return self._f(_lib.F_DZ_UNAPPLIED, dtype=np.float64)

    @property
    def qs_int_x(self):
        return self._f(_lib.F_QS_INT_X, self.state.grid.qx_shape)

    @property
    def qs_comp_x(self):
        return self._f(_lib.F_QS_COMP_X, self.state.grid.qx_shape)

    @property
    def qs_int_y(self):
        return self._f(_lib.F_QS_INT_Y, self.state.grid.qy_shape)

    @property
    def qs_comp_y(self):
        return self._f(_lib.F_QS_COMP_Y, self.state.grid.qy_shape)

    @property
    def inv_one_minus_p(self) -> float:
        return 1.0 / (1.0 - self.porosity)

    def bed_change_numpy(self) -> np.ndarray:
        return self.dz_cum.numpy()

    def totals(self) -> dict:
        """Sum, sum of magnitudes, min and max of dz_cum and the sum of dz_unapplied,
        reduced on the device (float64, fixed order): no field leaves the GPU."""
        import ctypes

        out = (ctypes.c_double * 5)()
        _lib.call("rbs_sed_totals", self.state._handle, out)
        return {"sum": out[0], "sum_abs": out[1], "min": out[2], "max": out[3],
                "unapplied": out[4]}

    def solid_volume(self, cell_area: float) -> float:
        return self.totals()["sum"] * cell_area * (1.0 - self.porosity)

    def banked_volume(self, cell_area: float) -> float:
        return self.totals()["unapplied"] * cell_area * (1.0 - self.porosity)

    def summary(self) -> str:
        grain = (f"d50 = {1000.0 * self.d50_min:.2f}..{1000.0 * self.d50_max:.2f} mm"
                 if self.d50_min != self.d50_max else f"d50 = {1000.0 * self.d50_max:.2f} mm")
        zeros = f", {self.zero_d50_cells} cells with no grain size" if self.zero_d50_cells else ""
        return f"{grain}, porosity {self.porosity:.2f}{zeros}"

    def as_attrs(self) -> dict:
        return {"d50_min_m": self.d50_min, "d50_max_m": self.d50_max, "porosity": self.porosity,
                "zero_d50_cells": self.zero_d50_cells}


def validate_grain_size(d50, grid) -> np.ndarray:
    """The grain-size field ``arm_sediment`` accepts (sediment.py:619-655): one value for
    every cell, or a field of exactly the grid's shape; finite and non-negative, where
    a cell of zero carries no sediment at all."""
    ny, nx = grid.shape
    if np.isscalar(d50):
        field = np.full((ny, nx), float(d50), dtype=np.float32)
    else:
        field = np.array(d50, dtype=np.float32, order="C", copy=True)
        if field.shape != (ny, nx):
            raise SedimentError(f"grain size d50 has shape {field.shape}, but the grid is "
                                f"{(ny, nx)}")
    if not np.isfinite(field).all():
        raise SedimentError("grain size d50 holds non-finite values")
    if (field < 0).any():
        raise SedimentError("grain size d50 must be >= 0 in every cell; 0 makes a cell inert")
    return field


def arm_sediment(state, d50, porosity: float) -> SedimentState:
    """Arm morphology on ``state`` (sediment.py:658-709).  The bed as it stands is the
    reference ``z0`` every later bed is rebuilt from; a state can be armed only once,
    since arming again would take ``z0`` from a bed that has already moved and lose the
    cumulative change that moved it."""
    if getattr(state, "sediment", None) is not None:
        raise SedimentError("this state is already armed for morphology (z0 was taken when it "
                            "was first armed)")
    if not 0.0 < porosity < 1.0:
        raise SedimentError(f"bed porosity must lie strictly between 0 and 1, got {porosity}")
    grain = validate_grain_size(d50, state.grid)
    state._arm(_lib.ARM_SED)
    sed = SedimentState(state, porosity, float(grain.min()), float(grain.max()),
                        int((grain == 0).sum()))
    sed.d50.assign(grain)
    state.sediment = sed
    return sed


# --- float64 host references of the law (sediment.py:719-801) --------------------------
def _f64(*vals):
    return tuple(np.asarray(v, dtype=np.float64) for v in vals)


def shields_from_flow(q, h, n, d50, radius=None) -> np.ndarray:
    """Shields number of discharge q over depth h (hydraulic radius h unless given)."""
    q, h, n, d50 = _f64(q, h, n, d50)
    r = h if radius is None else np.asarray(radius, dtype=np.float64)
    shear = GRAVITY * n * n * q * q / (h * h * np.cbrt(r))  # tau / rho
    return shear / (SUBMERGED_SG * GRAVITY * d50)


def capacity_from_flow(q, h, n, d50, radius=None) -> np.ndarray:
    """MPM bedload capacity per unit width (m^2/s), magnitude."""
    theta = shields_from_flow(q, h, n, d50, radius)
    (d50,) = _f64(d50)
    excess = np.maximum(theta - SHIELDS_CRITICAL, 0.0)
    return MPM_COEFFICIENT * excess**MPM_EXPONENT * np.sqrt(SUBMERGED_SG * GRAVITY * d50**3)


def bed_celerity(q, h, n, d50, porosity: float, width=None, dx=None) -> np.ndarray:
    """Migration speed of a low bed wave (m/s) under the MPM law at unit discharge q.

    Exner linearised about uniform flow: raising the bed by dz at fixed q thins the flow
    by dz, and theta ~ h^(-kappa) with kappa = 2 + (1/3) dln(R)/dln(h), so
    c_b = dq_s/dz / (1 - p) = 12 kappa (theta/h) sqrt(theta - theta_c) sqrt(s' g d50^3)
    / (1 - p).  Wide section: R = h, kappa = 7/3.  A sub-grid channel of ``width`` w
    uses R = w h / (w + 2h), kappa = 2 + w / (3 (w + 2h)), and with ``dx`` the result
    is scaled by the conveyance fraction w/dx (sediment.py:747-801)."""
    q, h, n, d50 = _f64(q, h, n, d50)
    if width is None:
        radius, kappa, share = h, 7.0 / 3.0, 1.0
    else:
        (w,) = _f64(width)
        radius = w * h / (w + 2.0 * h)
        kappa = 2.0 + w / (3.0 * (w + 2.0 * h))
        share = 1.0 if dx is None else w / float(dx)
    theta = shields_from_flow(q, h, n, d50, radius)
    excess = np.maximum(theta - SHIELDS_CRITICAL, 0.0)
    slope = (MPM_COEFFICIENT * MPM_EXPONENT * kappa * (theta / h) * np.sqrt(excess)
             * np.sqrt(SUBMERGED_SG * GRAVITY * d50**3))
    return share * slope / (1.0 - float(porosity))


def morphological_courant(celerity: float, interval_s: float, dx: float) -> float:
    """Cells a bed wave crosses in one slow interval, c_b interval / dx (sediment.py:811).
    Above one it measures the splitting of Exner from the water, not the physics."""
    return float(celerity) * float(interval_s) / float(dx)


# The relative submergence h_col/d50 from which the MPM law is read inside its range
# (sediment.py:845); it only selects the cells one reported reduction runs over.
MORPH_REGIME_FLOOR = 35.0


def column_depth_field(state) -> np.ndarray:
    """The depth the transport law sees, per cell (float64): h, or for a channelled cell
    the water column of the storage curve (channels.py:75-85)."""
    h = state.h.numpy().astype(np.float64)
    chan = state.channels
    if chan is None:
        return h
    dx = state.grid.dx
    w = chan.w.numpy().astype(np.float64)
    d = chan.d.numpy().astype(np.float64)
    channelled = (w > 0.0) & (d > 0.0)
    w1 = np.where(channelled, w, 1.0)
    bankfull = w1 * d / dx
    col = np.where(h <= bankfull, h * dx / w1, d + (h - bankfull))
    return np.where(channelled, col, h)


def submergence_field(state) -> np.ndarray:
    """h_col / d50 per cell, 0 where the cell has no grain size (sediment.py:870-880)."""
    sed = state.sediment
    if sed is None:
        raise SedimentError("submergence_field needs a state armed with arm_sediment()")
    d50 = sed.d50.numpy().astype(np.float64)
    has = d50 > 0.0
    return np.where(has, column_depth_field(state) / np.where(has, d50, 1.0), 0.0)


def _cell_flow(qa: np.ndarray, qb: np.ndarray) -> np.ndarray:
    """|q| at cell centres from the mean of each pair of opposite faces."""
    ux = 0.5 * (qa[:, :-1] + qa[:, 1:])
    uy = 0.5 * (qb[:-1, :] + qb[1:, :])
    return np.hypot(ux, uy)


def celerity_field(state) -> np.ndarray:
    """Bed-wave celerity (m/s) of every cell under the flow the state carries now
    (sediment.py:883-957): the wide-section celerity of the cell's centred discharge and
    depth, and in a channelled cell the larger of that and the channel component's
    celerity (its own discharge, column depth, roughness and conveyance fraction).  A
    dry cell, one with no grain size and one with no flow give exactly 0."""
    sed = state.sediment
    if sed is None:
        raise SedimentError("celerity_field needs a state armed with arm_sediment()")
    h = state.h.numpy().astype(np.float64)
    n = state.n.numpy().astype(np.float64)
    d50 = sed.d50.numpy().astype(np.float64)
    grain = d50 > 0.0
    d50_1 = np.where(grain, d50, 1.0)
    q = _cell_flow(state.qx.numpy().astype(np.float64), state.qy.numpy().astype(np.float64))
    live = grain & (h >= H_DRY) & (q > 0.0)
    c = np.where(live, bed_celerity(np.where(live, q, 0.0), np.where(live, h, 1.0), n, d50_1,
                                    sed.porosity), 0.0)
    chan = state.channels
    if chan is None:
        return c
    w = chan.w.numpy().astype(np.float64)
    d = chan.d.numpy().astype(np.float64)
    channelled = (w > 0.0) & (d > 0.0)
    col = column_depth_field(state)
    q_ch = _cell_flow(chan.qx_ch.numpy().astype(np.float64), chan.qy_ch.numpy().astype(np.float64))
    live_ch = channelled & grain & (col >= H_DRY) & (q_ch > 0.0)
    c_ch = bed_celerity(np.where(live_ch, q_ch, 0.0), np.where(live_ch, col, 1.0),
                        chan.n.numpy().astype(np.float64), d50_1, sed.porosity,
                        width=np.where(channelled, w, 1.0), dx=state.grid.dx)
    return np.maximum(np.where(live_ch, c_ch, 0.0), c)


@dataclass(frozen=True)
class CourantSummary:
    """One activation's morphological Courant number, reduced four ways
    (sediment.py:960-999): the field peak (what the run warns on), its peak over the
    cells whose bed moved in this activation, over the cells the law is read inside its
    range (h_col/d50 >= MORPH_REGIME_FLOOR), and the share of this activation's gross
    |dz| that fell where the number reached MORPH_COURANT_GATE."""

    celerity_m_s: float
    courant: float
    courant_moving: float
    courant_in_regime: float
    over_courant_share: float
    courant_cells: int
    live_cells: int

    def as_dict(self) -> dict:
        return dict(self.__dict__)


def courant_summary(state, dz: np.ndarray, interval_s: float) -> CourantSummary:
    """Reduce the celerity field against ``dz``, the bed change this activation applied
    (sediment.py:1002-1040)."""
    dx = state.grid.dx
    cel = celerity_field(state)
    peak = float(cel.max())
    per_cell = cel * (float(interval_s) / float(dx))
    moved = np.abs(np.asarray(dz, dtype=np.float64))
    gross = float(moved.sum())
    live = cel > 0.0
    changed = moved > 0.0
    regime = live & (submergence_field(state) >= MORPH_REGIME_FLOOR)
    over = per_cell >= MORPH_COURANT_GATE

    def peak_of(mask) -> float:
        return float(per_cell[mask].max()) if mask.any() else 0.0

    return CourantSummary(
        celerity_m_s=peak,
        courant=morphological_courant(peak, interval_s, dx),
        courant_moving=peak_of(changed),
        courant_in_regime=peak_of(regime),
        over_courant_share=float(moved[over].sum() / gross) if gross > 0.0 else 0.0,
        courant_cells=int(over.sum()),
        live_cells=int(live.sum()),
    )
