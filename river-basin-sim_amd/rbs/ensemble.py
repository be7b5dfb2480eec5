"""Batched ensembles: many local-inertial scenarios advanced by one launch per step.

A :class:`Batch` wraps ``rbs_batch_*``: members are states with their own device
run parameters (:class:`rbs.driver.DeviceRun`) and stop time; every ``advance`` runs
batched steps (one fused LI launch over all members' tiles + one finalize with a
block per member) until members reach their stop times.  Each member keeps its own
dt, clock and float64 ledger, and its arithmetic is exactly that of a solo run, so
results are bitwise those of the member advanced alone.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from rbs import _lib
from rbs.driver import DeviceRun


class Batch:
    def __init__(self, capacity: int):
        h = ctypes.c_void_p()
        _lib.call("rbs_batch_create", int(capacity), ctypes.byref(h))
        self._handle = h
        self.capacity = int(capacity)
        self.members: dict[int, DeviceRun] = {}

    def close(self) -> None:
        if getattr(self, "_handle", None):
            _lib.call("rbs_batch_destroy", self._handle)
            self._handle = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def add(self, run: DeviceRun, t_stop: float) -> int:
        """Admit a scenario; it steps from its current clock to ``t_stop``."""
        slot = ctypes.c_int32()
        _lib.call("rbs_batch_add", self._handle, run.state._handle, ctypes.byref(run.params),
                  float(t_stop), ctypes.byref(slot))
        self.members[int(slot.value)] = run
        return int(slot.value)

    def remove(self, slot: int) -> DeviceRun:
        _lib.call("rbs_batch_remove", self._handle, int(slot))
        return self.members.pop(int(slot))

    def advance(self, steps: int) -> int:
        """Run up to ``steps`` batched steps; returns the number of members still running."""
        n = ctypes.c_int32()
        _lib.call("rbs_batch_advance", self._handle, int(steps), ctypes.byref(n))
        return int(n.value)

    def set_variants(self, variant: int, variant_sed: int) -> None:
        """LI kernel variants for members without / with the bedload integral."""
        _lib.call("rbs_batch_set_variant", self._handle, int(variant), int(variant_sed))

    def variants(self) -> tuple[int, int]:
        v, vs = ctypes.c_int32(), ctypes.c_int32()
        _lib.call("rbs_batch_get_variant", self._handle, ctypes.byref(v), ctypes.byref(vs))
        return int(v.value), int(vs.value)

    def launch(self, steps: int) -> None:
        """Enqueue up to ``steps`` batched steps and return at once (see collect)."""
        _lib.call("rbs_batch_launch", self._handle, int(steps))

    def poll(self) -> bool:
        """True once the launched steps have finished (never blocks)."""
        d = ctypes.c_int32()
        _lib.call("rbs_batch_poll", self._handle, ctypes.byref(d))
        return bool(d.value)

    def collect(self) -> int:
        """Wait for the launched steps; returns the number of members still running."""
        n = ctypes.c_int32()
        _lib.call("rbs_batch_collect", self._handle, ctypes.byref(n))
        return int(n.value)

    def active(self, slot: int) -> bool:
        a = ctypes.c_int32()
        _lib.call("rbs_batch_member_active", self._handle, int(slot), ctypes.byref(a))
        return bool(a.value)

    def stats(self) -> dict:
        """Membership-change cost so far (table rebuilds + graph captures) and replays."""
        r, g = ctypes.c_int64(), ctypes.c_int64()
        ms, gpu = ctypes.c_double(), ctypes.c_double()
        _lib.call("rbs_batch_stats", self._handle, ctypes.byref(r), ctypes.byref(ms),
                  ctypes.byref(g), ctypes.byref(gpu))
        return {"rebuilds": int(r.value), "rebuild_s": ms.value / 1e3,
                "graph_launches": int(g.value), "gpu_s": gpu.value / 1e3}

    def run_to_completion(self, chunk: int = 64) -> None:
        while self.advance(chunk) > 0:
            pass


@dataclass
class EnsembleMember:
    """Inputs of one ensemble scenario (the seeded generator's output)."""

    bed: np.ndarray
    manning: np.ndarray | float
    rain: np.ndarray  # m/s
    dx: float
    seed: int


def ensemble_inputs(count: int, *, seed: int = 20261205, n_min: int = 256, n_max: int = 2048,
                    dx: float = 10.0) -> list[EnsembleMember]:
    """The batched-ensemble workload: sides log-uniform in [n_min, n_max], each on its own
    seeded fBm DEM, Manning 0.035 x U[0.7, 1.3], rain median U[5, 40] mm/h (lognormal
    field, sigma 0.6, 2 km correlation)."""
    from rbs import terrain

    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
        mult = rng.uniform(0.7, 1.3)
        median = rng.uniform(5.0, 40.0)
        s = seed + 1000 * (k + 1)
        bed = terrain.fbm_dem(n, n, hurst=0.8, relief=200.0, seed=s)
        rain = terrain.lognormal_field(n, n, median=median, sigma=0.6,
                                       corr_cells=2000.0 / dx, seed=s + 1) * np.float32(terrain.MM_H)
        out.append(EnsembleMember(bed=bed, manning=float(np.float32(0.035 * mult)),
                                  rain=rain.astype(np.float32), dx=dx, seed=s))
    return out


def make_member_state(m: EnsembleMember):
    """A device state for one ensemble member: rain field (Kahan), all edges open."""
    from rbs.core.state import State

    from rbs.autotune import tune

    st = State.from_bed(m.bed, m.dx, depth=0.0, manning=m.manning)
    st.set_rain_field(m.rain)
    st.set_open_boundaries({e: "open" for e in ("north", "south", "east", "west")})
    st.arm_source_compensation()
    tune(st)  # a scenario run on its own uses the fastest kernel for its shape
    return st
