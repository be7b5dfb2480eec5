"""Solver state on the GPU (reference solver/core/state.py:38-256).

``State`` keeps the reference's field names, arming rules and error behaviour;
the fields live in device memory owned by ``librbs`` (one handle per state) and
each is exposed as a :class:`DeviceField` with ``numpy()`` / ``assign()``, the two
operations the reference's callers use on its Warp arrays.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from rbs import _lib
from rbs.core.grid import H_DRY, Grid

_EDGES = ("north", "south", "east", "west")  # state.py:35


@dataclass
class DeviceField:
    """A device array of the state, transferred dense in the reference's layout."""

    state: State
    field_id: int
    shape: tuple[int, int]
    dtype: type = np.float32

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        _lib.call("rbs_download", self.state._handle, self.field_id,
                  out.ctypes.data_as(ctypes.c_void_p), out.nbytes)
        return out

    def assign(self, values) -> None:
        arr = values.numpy() if hasattr(values, "numpy") else values
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        if arr.shape != self.shape:
            raise ValueError(f"shape {arr.shape} != {self.shape}")
        _lib.call("rbs_upload", self.state._handle, self.field_id,
                  arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes)

    def zero_(self) -> None:
        self.assign(np.zeros(self.shape, self.dtype))


def _device_ok(device: str | None) -> str:
    if device in (None, "cuda", "cuda:0", "gpu", "hip", "hip:0"):
        return "cuda:0"
    raise ValueError(
        f"device {device!r}: this solver runs only on the GPU through librbs (no CPU path)"
    )


class State:
    """Mutable float32 fields of one simulation, resident on the GPU."""

    def __init__(self, grid: Grid, device: str = "cuda:0", arm: int = 0):
        self.grid = grid
        self.device = _device_ok(device)
        h = ctypes.c_void_p()
        _lib.call("rbs_state_create", grid.ny, grid.nx, float(grid.dx), arm, ctypes.byref(h))
        self._handle = h
        self.boundaries: dict[str, str] = {e: "closed" for e in _EDGES}
        self.open_edges: frozenset[str] = frozenset()
        self.stage_curves: dict[str, list[tuple[float, float]]] = {}
        self.channels = None  # rbs.core.channels.ChannelState once armed
        self.sediment = None  # rbs.core.sediment.SedimentState once armed
        self.hllc = None

    # --- lifecycle -------------------------------------------------------------
    def close(self) -> None:
        if getattr(self, "_handle", None):
            try:
                _lib.call("rbs_state_destroy", self._handle)
            finally:
                self._handle = None

    def __del__(self):  # pragma: no cover - interpreter teardown order varies
        try:
            self.close()
        except Exception:
            pass

    def _armed(self) -> int:
        v = ctypes.c_uint32()
        _lib.call("rbs_armed", self._handle, ctypes.byref(v))
        return int(v.value)

    def _arm(self, flags: int) -> None:
        _lib.call("rbs_arm", self._handle, flags)

    def device_bytes(self) -> int:
        v = ctypes.c_uint64()
        _lib.call("rbs_state_bytes", self._handle, ctypes.byref(v))
        return int(v.value)

    # --- fields ----------------------------------------------------------------
    def _cell(self, fid: int) -> DeviceField:
        return DeviceField(self, fid, self.grid.shape)

    @property
    def h(self) -> DeviceField:
        return self._cell(_lib.F_H)

    @h.setter
    def h(self, v) -> None:
        self.h.assign(v)

    @property
    def z(self) -> DeviceField:
        return self._cell(_lib.F_Z)

    @z.setter
    def z(self, v) -> None:
        self.z.assign(v)

    @property
    def n(self) -> DeviceField:
        return self._cell(_lib.F_N)

    @n.setter
    def n(self, v) -> None:
        self.n.assign(v)

    @property
    def qx(self) -> DeviceField:
        return DeviceField(self, _lib.F_QX, self.grid.qx_shape)

    @qx.setter
    def qx(self, v) -> None:
        self.qx.assign(v)

    @property
    def qy(self) -> DeviceField:
        return DeviceField(self, _lib.F_QY, self.grid.qy_shape)

    @qy.setter
    def qy(self, v) -> None:
        self.qy.assign(v)

    @property
    def eta(self) -> DeviceField:
        return self._cell(_lib.F_ETA)

    def _optional(self, flag: int, fid: int) -> DeviceField | None:
        return self._cell(fid) if self._armed() & flag else None

    @property
    def hu(self) -> DeviceField | None:
        return self._optional(_lib.ARM_HLLC, _lib.F_HU)

    @property
    def hv(self) -> DeviceField | None:
        return self._optional(_lib.ARM_HLLC, _lib.F_HV)

    @property
    def h_comp(self) -> DeviceField | None:
        return self._optional(_lib.ARM_COMP, _lib.F_H_COMP)

    @property
    def rain(self) -> DeviceField | None:
        return self._optional(_lib.ARM_RAIN, _lib.F_RAIN)

    @property
    def infil(self) -> DeviceField | None:
        return self._optional(_lib.ARM_INFIL, _lib.F_INFIL)

    # --- arming (state.py:98-171) -------------------------------------------------
    def set_infiltration(self, infil: np.ndarray) -> None:
        """Attach an infiltration-rate field (m/s); the device ledger banks the loss."""
        if np.shape(infil) != self.grid.shape:
            raise ValueError(f"infiltration shape {np.shape(infil)} != grid {self.grid.shape}")
        self._arm(_lib.ARM_INFIL)
        self._cell(_lib.F_INFIL).assign(infil)

    def set_rain_field(self, rain: np.ndarray) -> None:
        """Attach a spatial rainfall-rate field (m/s); scaled on/off over time."""
        if np.shape(rain) != self.grid.shape:
            raise ValueError(f"rain-field shape {np.shape(rain)} != grid {self.grid.shape}")
        self._arm(_lib.ARM_RAIN)
        self._cell(_lib.F_RAIN).assign(rain)

    def set_open_boundaries(
        self,
        boundaries: dict[str, str],
        stage_curves: dict[str, list[tuple[float, float]]] | None = None,
    ) -> None:
        """Per-edge map closed|open|fixed_stage, plus curves for the stage edges."""
        bnd = {e: boundaries.get(e, "closed") for e in _EDGES}
        for e, v in bnd.items():
            if v not in _lib.BC:
                raise ValueError(f"unknown boundary type {v!r} on edge {e}")
        curves = {e: list(c) for e, c in (stage_curves or {}).items()
                  if bnd.get(e) == "fixed_stage"}
        missing = [e for e, v in bnd.items() if v == "fixed_stage" and e not in curves]
        if missing:
            raise ValueError(f"fixed_stage edges without a stage curve: {sorted(missing)}")
        types = (ctypes.c_int32 * 4)(*[_lib.BC[bnd[e]] for e in ("west", "east", "north", "south")])
        _lib.call("rbs_set_boundaries", self._handle, types)
        for e, c in curves.items():
            t = np.ascontiguousarray([p[0] for p in c], dtype=np.float64)
            lv = np.ascontiguousarray([p[1] for p in c], dtype=np.float64)
            _lib.call("rbs_set_stage_curve", self._handle, _lib.EDGE[e], len(t),
                      t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                      lv.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        self.boundaries = bnd
        self.open_edges = frozenset(e for e, v in bnd.items() if v == "open")
        self.stage_curves = curves

    def arm_source_compensation(self) -> None:
        """Arm per-cell Kahan compensation for areal sources (idempotent)."""
        self._arm(_lib.ARM_COMP)

    def arm_loss_accumulator(self) -> None:
        """Sinks bank into the device ledger, which always exists (no-op)."""

    def loss_volume(self, cell_area: float | None = None) -> float:
        """Cumulative sink volume (m^3): infiltration, edge drains and the banked
        rounding of reservoir deliveries, float64 on device."""
        led = self.ledger()
        area_scale = 1.0 if cell_area is None else cell_area / self.grid.cell_area
        return (led.infil_out + led.edge_out + led.release_out) * area_scale

    def ledger(self) -> _lib.Ledger:
        out = _lib.Ledger()
        _lib.call("rbs_ledger_read", self._handle, ctypes.byref(out))
        return out

    # --- construction (state.py:179-226) ---------------------------------------------
    @classmethod
    def from_bed(
        cls,
        bed: np.ndarray,
        dx: float,
        *,
        depth: np.ndarray | float = 0.0,
        manning: np.ndarray | float = 0.035,
        device: str = "cuda:0",
    ) -> State:
        """A state over ``bed`` with uniform or per-cell initial depth and roughness
        (state.py:179-226); every face discharge starts at zero."""
        z = np.ascontiguousarray(bed, dtype=np.float32)
        shape = z.shape

        def cells(name, value):
            if np.isscalar(value):
                return np.full(shape, float(value), dtype=np.float32)
            arr = np.ascontiguousarray(value, dtype=np.float32)
            if arr.shape != shape:
                raise ValueError(f"{name} has shape {arr.shape}; the bed is {shape}")
            return arr

        h0, n0 = cells("depth", depth), cells("manning", manning)
        st = cls(Grid(ny=shape[0], nx=shape[1], dx=float(dx)), device=device)
        for dst, values in ((st.z, z), (st.n, n0), (st.h, h0)):
            dst.assign(values)
        return st

    def depth_numpy(self) -> np.ndarray:
        return self.h.numpy()

    def velocities_numpy(self) -> tuple[np.ndarray, np.ndarray]:
        """Cell-centred (u, v), 0 where h < H_DRY (state.py:232-256)."""
        h = self.h.numpy()
        wet = h >= H_DRY
        u = np.zeros_like(h)
        v = np.zeros_like(h)
        if self._armed() & _lib.ARM_HLLC:
            np.divide(self.hu.numpy(), h, out=u, where=wet)
            np.divide(self.hv.numpy(), h, out=v, where=wet)
        else:
            qx = self.qx.numpy()
            qy = self.qy.numpy()
            qx_c = 0.5 * (qx[:, :-1] + qx[:, 1:])
            qy_c = 0.5 * (qy[:-1, :] + qy[1:, :])
            np.divide(qx_c, h, out=u, where=wet)
            np.divide(qy_c, h, out=v, where=wet)
        return u.astype(np.float32), v.astype(np.float32)

    def sync(self) -> None:
        _lib.call("rbs_sync", self._handle)
