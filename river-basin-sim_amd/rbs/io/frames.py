"""Output frames formed on the device and copied to the host off the step stream.

The reference's output step (run.py:389-400) reads ``h``, ``velocities_numpy()``
(state.py:232-256) and the bed change back synchronously at each output time.  A
:class:`FrameQueue` replaces those reads: ``capture()`` enqueues one kernel after the
steps already on the state's stream, which forms depth, cell-centred u, v and the
float32 bed change into a device slot with the frame's clock, and a copy stream
moves the slot to pinned host memory while the state goes on stepping.  ``wait``
hands back numpy views of the pinned planes (valid until ``release``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from rbs import _lib

_FIELD_BITS = {"depth": _lib.FRAME_DEPTH, "u": _lib.FRAME_U, "v": _lib.FRAME_V,
               "bed_change": _lib.FRAME_BED_CHANGE}


@dataclass
class Frame:
    """One output frame: the device clock and dense (ny, nx) float32 planes."""

    slot: int
    time: float
    depth: np.ndarray | None
    u: np.ndarray | None
    v: np.ndarray | None
    bed_change: np.ndarray | None


class FrameQueue:
    """``slots`` frames of ``state`` in flight at once (one host thread per queue)."""

    def __init__(self, state, fields=("depth", "u", "v"), slots: int = 2):
        unknown = set(fields) - set(_FIELD_BITS)
        if unknown:
            raise ValueError(f"unknown frame fields {sorted(unknown)}")
        self.state = state
        self.fields = tuple(fields)
        mask = 0
        for f in fields:
            mask |= _FIELD_BITS[f]
        h = ctypes.c_void_p()
        _lib.call("rbs_frames_create", state._handle, mask, int(slots), ctypes.byref(h))
        self._handle = h
        self.slots = int(slots)

    def close(self) -> None:
        if getattr(self, "_handle", None):
            _lib.call("rbs_frames_destroy", self._handle)
            self._handle = None

    def __del__(self):  # pragma: no cover - interpreter teardown order varies
        try:
            self.close()
        except Exception:
            pass

    def capture(self) -> int:
        """Enqueue a frame of the state as it stands after its enqueued steps."""
        slot = ctypes.c_int32()
        _lib.call("rbs_frames_capture", self._handle, ctypes.byref(slot))
        return int(slot.value)

    def ready(self, slot: int) -> bool:
        r = ctypes.c_int32()
        _lib.call("rbs_frames_ready", self._handle, int(slot), ctypes.byref(r))
        return bool(r.value)

    def wait(self, slot: int) -> Frame:
        """Block until the slot is on the host; views of its pinned planes."""
        t = ctypes.c_double()
        ptrs = [ctypes.POINTER(ctypes.c_float)() for _ in range(4)]
        _lib.call("rbs_frames_wait", self._handle, int(slot), ctypes.byref(t),
                  *[ctypes.byref(p) for p in ptrs])
        shape = self.state.grid.shape
        views = [np.ctypeslib.as_array(p, shape=shape) if p else None for p in ptrs]
        return Frame(int(slot), float(t.value), *views)

    def release(self, slot: int) -> None:
        _lib.call("rbs_frames_release", self._handle, int(slot))

    def device_bytes(self) -> int:
        v = ctypes.c_uint64()
        _lib.call("rbs_frames_bytes", self._handle, ctypes.byref(v))
        return int(v.value)
