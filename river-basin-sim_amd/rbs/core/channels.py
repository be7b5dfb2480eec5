"""Sub-grid channels: geometry, storage curve and arming (reference
solver/core/channels.py).

A cell may carry a channel narrower than itself: width ``w`` (``0 < w <= dx``),
bank-full depth ``d`` below the floodplain bed ``z`` and its own roughness ``n_ch``.
The state variable stays the cell-mean depth ``h`` (volume per plan area), so
continuity, the limiter and the ledger are unchanged; the channel enters through
the storage curve ``h -> eta`` (channels.py:59-72)::

    h_bf = w d / dx
    h <= h_bf :  eta = z - d + h dx / w
    h >  h_bf :  eta = z + (h - h_bf)

and through the two-component faces, which the fused channel kernel of librbs
evaluates (local_inertial.py:266-500, sediment.py:276-377).  Arming switches the
state's local-inertial step, its CFL reduction (over the water column) and its
bedload integral onto the sub-grid forms; ``w = 0`` everywhere reproduces the plain
step bit for bit.

This module holds the host side: geometry validation (channels.py:214-286), the
float64 host storage curve and its inverse (:141-175), and ``arm_channels``
(:289-330), which uploads the validated fields to the device.
"""

from __future__ import annotations

import warnings

import numpy as np

from rbs import _lib
from rbs.core.grid import Grid

# channels.py:51: a channel narrower than dx/1000 collapses the timestep.
WIDTH_WARN_RATIO = 1000.0


class ChannelError(ValueError):
    """Sub-grid channel geometry is inconsistent with the grid or with itself."""


def eta_from_h(h, z, w, d, dx: float) -> np.ndarray:
    """Host storage curve in float64 (channels.py:141-158)."""
    h, z, w, d = (np.asarray(v, dtype=np.float64) for v in (h, z, w, d))
    has = (w > 0.0) & (d > 0.0)
    safe_w = np.where(has, w, 1.0)
    h_bf = np.where(has, w * d / dx, 0.0)
    in_bank = np.where(has, z - d + h * dx / safe_w, 0.0)
    over_bank = z + (h - h_bf)
    return np.where(has, np.where(h <= h_bf, in_bank, over_bank), z + h)


def h_from_eta(eta, z, w, d, dx: float) -> np.ndarray:
    """Inverse storage curve: the cell-mean depth holding surface ``eta`` (:161-175)."""
    eta, z, w, d = (np.asarray(v, dtype=np.float64) for v in (eta, z, w, d))
    has = (w > 0.0) & (d > 0.0)
    h_bf = np.where(has, w * d / dx, 0.0)
    channel = (eta - (z - d)) * w / dx
    depth = np.where(has, np.where(eta <= z, channel, h_bf + (eta - z)), eta - z)
    return np.clip(depth, 0.0, None)


class ChannelState:
    """Device-side channel geometry and split face discharges of one state
    (channels.py:178-211).  ``qx_ch`` is discharge per unit channel width, ``qx_fp``
    per unit floodplain width; the kernel recombines them into ``qx``."""

    def __init__(self, state, width_max: float, depth_max: float, channel_cells: int):
        self.state = state
        self.width_max = float(width_max)
        self.depth_max = float(depth_max)
        self.channel_cells = int(channel_cells)

    def _f(self, fid, shape=None):
        from rbs.core.state import DeviceField

        return DeviceField(self.state, fid, shape or self.state.grid.shape)

    @property
    def w(self):
        return self._f(_lib.F_CH_W)

    @property
    def d(self):
        return self._f(_lib.F_CH_D)

    @property
    def n(self):
        return self._f(_lib.F_CH_N)

    @property
    def qx_ch(self):
        return self._f(_lib.F_QX_CH, self.state.grid.qx_shape)

    @property
    def qx_fp(self):
        return self._f(_lib.F_QX_FP, self.state.grid.qx_shape)

    @property
    def qy_ch(self):
        return self._f(_lib.F_QY_CH, self.state.grid.qy_shape)

    @property
    def qy_fp(self):
        return self._f(_lib.F_QY_FP, self.state.grid.qy_shape)

    def summary(self) -> str:
        return (f"sub-grid channels in {self.channel_cells} cells (widths up to "
                f"{self.width_max:.1f} m, bank-full depths up to {self.depth_max:.2f} m)")

    def as_attrs(self) -> dict:
        # the result store's attribute keys for the channel geometry (channels.py:205-217)
        keys = ("channel_cells", "width_max_m", "depth_max_m")
        return dict(zip(keys, (self.channel_cells, self.width_max, self.depth_max)))


def _channel_field(name: str, values, shape) -> np.ndarray:
    if np.shape(values) != shape:
        raise ChannelError(f"the channel {name} field has shape {np.shape(values)}; the grid "
                           f"is {shape}")
    return np.array(values, dtype=np.float32, order="C", copy=True)


def validate_geometry(width, depth, manning, grid: Grid):
    """The channel geometry ``arm_channels`` accepts (channels.py:214-286), as float32
    copies ``(w, d, n)``; a cell without a channel comes back as w = d = 0.

    Refused: fields of another shape than the grid, non-finite or negative widths and
    depths, a channel wider than its cell (a sub-grid channel lives inside one), a
    width without a depth or the reverse, and a channel cell without a positive
    roughness.  A channel far narrower than its cell is legal but sets the column depth
    dx/w and with it the timestep, so it is warned about."""
    shape = grid.shape
    w = _channel_field("width", width, shape)
    d = _channel_field("depth", depth, shape)
    n = _channel_field("manning", manning, shape)
    if not np.isfinite(w).all() or not np.isfinite(d).all():
        raise ChannelError("channel width or depth holds non-finite values")
    if min(float(w.min()), float(d.min())) < 0.0:
        raise ChannelError("channel width and depth must be >= 0 in every cell")
    too_wide = w > grid.dx  # float32 widths against the float64 cell size
    if too_wide.any():
        raise ChannelError(
            f"a channel wider than the cell in {int(too_wide.sum())} cells (up to "
            f"{float(w.max()):.1f} m against dx = {grid.dx:.1f} m): a sub-grid channel must "
            "fit inside its cell, so resolve this river on a finer grid")
    half = (w > 0) ^ (d > 0)
    if half.any():
        raise ChannelError(
            f"{int(half.sum())} cells give a channel width without a bank-full depth, or a "
            "depth without a width; a channel has both, and no channel is width = depth = 0")
    present = (w > 0) & (d > 0)
    if present.any() and not (n[present] > 0).all():
        raise ChannelError("every channel cell needs its own manning > 0")
    w = np.where(present, w, np.float32(0.0))
    d = np.where(present, d, np.float32(0.0))
    if present.any():
        w_min = float(w[present].min())
        if w_min < grid.dx / WIDTH_WARN_RATIO:
            warnings.warn(
                f"a {w_min:g} m channel in {grid.dx:g} m cells: the column depth grows as "
                "dx/w, and the timestep shrinks with it, so this run will be slow",
                stacklevel=2)
    return w, d, n


def arm_channels(state, width, depth, manning=None) -> ChannelState:
    """Attach validated channel geometry to ``state`` (channels.py:289-330).

    ``manning`` defaults to the floodplain roughness ``state.n``.  The face
    components start at zero.  Arming twice replaces the geometry and keeps the
    components.
    """
    fp_n = state.n.numpy()
    w, d, n = validate_geometry(
        np.asarray(width, dtype=np.float32),
        np.asarray(depth, dtype=np.float32),
        fp_n if manning is None else np.asarray(manning, dtype=np.float32),
        state.grid,
    )
    state._arm(_lib.ARM_CHAN)
    chan = ChannelState(state, float(w.max()), float(d.max()), int((w > 0).sum()))
    chan.w.assign(w)
    chan.d.assign(d)
    chan.n.assign(n)
    state.channels = chan
    return chan
