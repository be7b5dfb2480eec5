"""Reservoirs: barrier geometry plus a release rule on the slow clock (reference
solver/processes/reservoir.py and the ``Structure`` of solver/io/config.py:124-255).

* :func:`apply_barriers` raises the bed to each structure's crest at its cells, once
  before stepping; impoundment and overtopping are then ordinary shallow-water
  physics.
* :class:`ReservoirOperator` applies a dam's rule at each activation of its slow
  process: ``rbs_reservoir_release`` reads the gauge stage, evaluates the rule and
  moves ``min(Q dt_slow, pool volume)`` from the pool box to the outlet box in one
  device launch, banking the float32 rounding of the delivery as an accounted sink,
  so the transfer is mass-exact in the ledger.

The release changes depths between steps, and the reference's next step takes its
CFL dt from the changed field (``compute_dt`` opens the tick).  A device-driven run
therefore stops on the host at each activation, as it does for morphology; the
transfer itself is a single launch.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from rbs import _lib
from rbs.scheduler import SlowProcess

_STRUCTURE_KINDS = {"dam", "levee"}  # config.py:124
_RELEASE_RULES = {"none", "fixed", "target_stage"}  # :125
_DEFAULT_RELEASE_INTERVAL_S = 900.0  # :129


@dataclass
class Structure:
    """A dam or levee with an optional release rule (config.py:147-255): ``cells`` and
    ``crest_m`` are barrier geometry; a dam's rule moves water from the inclusive
    ``pool`` box to the ``outlet`` box every ``interval_s`` simulated seconds."""

    name: str = "dam"
    kind: str = "dam"
    cells: list = field(default_factory=list)
    crest_m: float = 0.0
    release_rule: str = "none"
    release_m3_s: float = 0.0
    target_stage_m: float | None = None
    release_max_m3_s: float = 0.0
    pool: tuple | None = None
    outlet: tuple | None = None
    interval_s: float = _DEFAULT_RELEASE_INTERVAL_S

    def __post_init__(self) -> None:
        n = self.name
        if self.kind not in _STRUCTURE_KINDS:
            raise ValueError(f"structure '{n}': type must be one of {sorted(_STRUCTURE_KINDS)}")
        if not self.cells:
            raise ValueError(f"structure '{n}': needs at least one barrier cell")
        if self.release_rule not in _RELEASE_RULES:
            raise ValueError(f"structure '{n}': release_rule must be one of {sorted(_RELEASE_RULES)}")
        if self.kind == "levee" and self.release_rule != "none":
            raise ValueError(f"structure '{n}': a levee is barrier geometry only; use type='dam' "
                             "for a release rule")
        if self.interval_s <= 0:
            raise ValueError(f"structure '{n}': interval_s must be > 0")
        if self.release_rule == "none":
            return
        if self.pool is None or self.outlet is None:
            raise ValueError(f"structure '{n}': release_rule='{self.release_rule}' needs both a "
                             "'pool' box and an 'outlet' cell (where the released water is delivered)")
        if len(self.outlet) == 2:  # a single cell is a 1x1 box
            self.outlet = (self.outlet[0], self.outlet[1], self.outlet[0], self.outlet[1])
        r0, c0, r1, c1 = self.pool
        if r1 < r0 or c1 < c0:
            raise ValueError(f"structure '{n}': pool must be [row0, col0, row1, col1]")
        o0, p0, o1, p1 = self.outlet
        if o1 < o0 or p1 < p0:
            raise ValueError(f"structure '{n}': outlet must be [row, col] or [row0, col0, row1, col1]")
        if not (o1 < r0 or o0 > r1 or p1 < c0 or p0 > c1):
            raise ValueError(f"structure '{n}': the outlet {self.outlet} overlaps the pool "
                             f"{self.pool}; the release would just shuffle water within the reservoir")
        if self.release_rule == "fixed" and self.release_m3_s <= 0:
            raise ValueError(f"structure '{n}': release_rule='fixed' needs release_m3_s > 0")
        if self.release_rule == "target_stage":
            if self.target_stage_m is None:
                raise ValueError(f"structure '{n}': release_rule='target_stage' needs target_stage_m")
            if self.release_max_m3_s <= 0:
                raise ValueError(f"structure '{n}': release_rule='target_stage' needs "
                                 "release_max_m3_s > 0 (the cap the proportional rule scales toward)")
            if self.target_stage_m >= self.crest_m:
                raise ValueError(f"structure '{n}': target_stage_m ({self.target_stage_m}) must be "
                                 f"below crest_m ({self.crest_m}) -- the rule ramps between the two")

    def discharge_at(self, stage: float | None) -> float:
        """The release (m^3/s) the rule asks for at a pool stage (config.py:239-255)."""
        if self.release_rule == "none" or stage is None:
            return 0.0
        if self.release_rule == "fixed":
            return float(self.release_m3_s)
        span = self.crest_m - float(self.target_stage_m)
        frac = (stage - float(self.target_stage_m)) / span
        return float(self.release_max_m3_s) * min(max(frac, 0.0), 1.0)


def apply_barriers(bed: np.ndarray, structures: list[Structure]) -> np.ndarray:
    """Raise the bed to each crest at the structure's cells, never lowering it
    (reservoir.py:60-80); returns a new array (the input itself when there are none)."""
    if not structures:
        return bed
    out = np.array(bed, dtype=np.float32, copy=True)
    ny, nx = out.shape
    for s in structures:
        for i, j in s.cells:
            if not (0 <= i < ny and 0 <= j < nx):
                raise ValueError(f"structure '{s.name}': cell {(i, j)} is outside the {ny}x{nx} grid")
            out[i, j] = max(float(out[i, j]), float(s.crest_m))
    return out


@dataclass
class ReleaseRecord:
    """One activation (reservoir.py:133-147)."""

    time: float
    stage: float | None
    discharge_m3_s: float
    volume_m3: float

    def as_dict(self) -> dict:
        return {"time": self.time, "stage": self.stage, "discharge_m3_s": self.discharge_m3_s,
                "volume_m3": self.volume_m3}


def check_boxes(structure: Structure, shape: tuple[int, int]) -> None:
    """The pool and outlet must lie inside the grid (reservoir.py:166-177)."""
    ny, nx = shape
    r0, c0, r1, c1 = structure.pool
    if not (0 <= r0 <= r1 < ny and 0 <= c0 <= c1 < nx):
        raise ValueError(f"structure '{structure.name}': pool {structure.pool} is outside the "
                         f"{ny}x{nx} grid")
    o0, p0, o1, p1 = structure.outlet
    if not (0 <= o0 <= o1 < ny and 0 <= p0 <= p1 < nx):
        raise ValueError(f"structure '{structure.name}': outlet {structure.outlet} is outside the "
                         f"{ny}x{nx} grid")


def gauge_cell(structure: Structure, bed: np.ndarray) -> tuple[int, int]:
    """The stilling well: the pool box's lowest-bed cell (reservoir.py:181-184, 198-217),
    the first in row-major order on ties."""
    r0, c0, r1, c1 = structure.pool
    box = np.asarray(bed)[r0:r1 + 1, c0:c1 + 1]
    a, b = np.unravel_index(int(np.argmin(box)), box.shape)
    return r0 + int(a), c0 + int(b)


class ReservoirOperator:
    """A dam's release rule on the slow clock (reservoir.py:150-275)."""

    def __init__(self, structure: Structure, state):
        if structure.release_rule == "none":
            raise ValueError(f"structure '{structure.name}': has no release rule to operate")
        check_boxes(structure, state.grid.shape)
        self.structure = structure
        self.state = state
        self.records: list[ReleaseRecord] = []
        self._gauge = gauge_cell(structure, state.z.numpy())
        r = _lib.Reservoir()
        for k in range(4):
            r.pool[k] = int(structure.pool[k])
            r.outlet[k] = int(structure.outlet[k])
        r.gauge_row, r.gauge_col = self._gauge
        r.rule = _lib.RULE_FIXED if structure.release_rule == "fixed" else _lib.RULE_TARGET_STAGE
        r.release_m3_s = float(structure.release_m3_s)
        r.target_stage_m = float(structure.target_stage_m or 0.0)
        r.crest_m = float(structure.crest_m)
        r.release_max_m3_s = float(structure.release_max_m3_s)
        self._params = r

    def as_slow_process(self) -> SlowProcess:
        return SlowProcess(name=f"reservoir:{self.structure.name}",
                           interval=self.structure.interval_s, advance=self.advance)

    def pool_stage(self) -> float | None:
        """z + h at the gauge cell, ``None`` when it is dry (reservoir.py:198-217)."""
        gi, gj = self._gauge
        h = float(self.state.h.numpy()[gi, gj])
        if h <= 1.0e-3:
            return None
        return float(self.state.z.numpy()[gi, gj]) + h

    def advance(self, t: float, dt_slow: float) -> ReleaseRecord:
        """One activation: the rule's release over ``dt_slow`` seconds (reservoir.py:219-234)."""
        out = _lib.Release()
        _lib.call("rbs_reservoir_release", self.state._handle, ctypes.byref(self._params),
                  float(t), float(dt_slow), ctypes.byref(out))
        stage = None if math.isnan(out.stage) else float(out.stage)
        rec = ReleaseRecord(time=float(out.time), stage=stage,
                            discharge_m3_s=float(out.discharge_m3_s), volume_m3=float(out.volume_m3))
        self.records.append(rec)
        return rec

    @property
    def series(self) -> list[dict]:
        return [r.as_dict() for r in self.records]


def build_operators(structures: list[Structure], state) -> list[ReservoirOperator]:
    """An operator for every structure with a release rule (reservoir.py:278-280)."""
    return [ReservoirOperator(s, state) for s in structures if s.release_rule != "none"]
