"""Bedload transport (Meyer-Peter-Mueller) and the Exner bed update
(reference solver/core/sediment.py).

The transport integral is accumulated inside the fused local-inertial kernel
(after the limiter, before continuity, exactly where the reference places it,
local_inertial.py:680-687); the Exner update and bed rebuild run at morphology
activations (``rbs_exner_activate``).  This module arms the state, validates
inputs and keeps the reference's float64 host references for the law.
"""

from __future__ import annotations

import numpy as np

from rbs import _lib
from rbs.core.grid import GRAVITY

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
    """sediment.py:619-655: broadcast a scalar, refuse negative / non-finite."""
    ny, nx = grid.shape
    if np.isscalar(d50):
        arr = np.full((ny, nx), float(d50), dtype=np.float32)
    else:
        arr = np.ascontiguousarray(d50, dtype=np.float32).copy()
        if arr.shape != (ny, nx):
            raise SedimentError(f"d50 shape {arr.shape} != grid {(ny, nx)}")
    if not np.isfinite(arr).all():
        raise SedimentError("d50 field contains non-finite values")
    if (arr < 0).any():
        raise SedimentError("d50 must be >= 0 everywhere (0 = an inert cell)")
    return arr


def arm_sediment(state, d50, porosity: float) -> SedimentState:
    """Arm morphology on ``state`` (sediment.py:658-709): z0 is captured from the bed
    as it is now; arming twice raises."""
    if getattr(state, "sediment", None) is not None:
        raise SedimentError(
            "state is already armed for morphology; re-arming would recapture z0 from "
            "a moved bed and discard the cumulative change that moved it")
    if not 0.0 < porosity < 1.0:
        raise SedimentError(f"porosity must be in (0, 1), got {porosity}")
    grain = validate_grain_size(d50, state.grid)
    state._arm(_lib.ARM_SED)
    sed = SedimentState(state, porosity, float(grain.min()), float(grain.max()),
                        int((grain == 0).sum()))
    sed.d50.assign(grain)
    state.sediment = sed
    return sed


# --- float64 host references of the law (sediment.py:719-801) --------------------------
def shields_from_flow(q, h, n, d50, radius=None) -> np.ndarray:
    q, h, n, d50 = (np.asarray(v, dtype=np.float64) for v in (q, h, n, d50))
    r = h if radius is None else np.asarray(radius, dtype=np.float64)
    tau_over_rho = GRAVITY * n**2 * q**2 / (h**2 * np.cbrt(r))
    return tau_over_rho / (SUBMERGED_SG * GRAVITY * d50)


def capacity_from_flow(q, h, n, d50, radius=None) -> np.ndarray:
    theta = shields_from_flow(q, h, n, d50, radius)
    d50 = np.asarray(d50, dtype=np.float64)
    excess = np.maximum(theta - SHIELDS_CRITICAL, 0.0)
    return MPM_COEFFICIENT * excess**MPM_EXPONENT * np.sqrt(SUBMERGED_SG * GRAVITY * d50**3)


def bed_celerity(q, h, n, d50, porosity: float) -> np.ndarray:
    """Linearised Exner celerity of a low bed wave for the wide-section law."""
    q, h, n, d50 = (np.asarray(v, dtype=np.float64) for v in (q, h, n, d50))
    theta = shields_from_flow(q, h, n, d50)
    excess = np.maximum(theta - SHIELDS_CRITICAL, 0.0)
    kappa = 7.0 / 3.0
    dqs_dz = (MPM_COEFFICIENT * MPM_EXPONENT * kappa * (theta / h) * np.sqrt(excess)
              * np.sqrt(SUBMERGED_SG * GRAVITY * d50**3))
    return dqs_dz / (1.0 - float(porosity))
