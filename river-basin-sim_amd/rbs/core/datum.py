"""Vertical datum shift (reference solver/core/datum.py:35-62, applied by
``shift_for_datum`` in solver/run.py:112-137).

Fields are float32, and the schemes work on the water surface eta = h + z, which
float32 resolves only to about ULP(z): 1e-6 m at a 10 m datum, 5e-4 m at 5000 m.  The
well-balanced cancellation of a lake at rest happens in exactly that arithmetic, so a
high datum shows up first as a lake that will not stay flat.  Moving the origin to a
reference elevation z_ref before stepping restores the precision; every absolute
elevation the run carries (the bed, fixed-stage levels, structure crests and target
stages) must move with it, and outputs add z_ref back to the bed.
"""

from __future__ import annotations

import dataclasses
from dataclasses import dataclass

import numpy as np


def resolve_datum(datum, bed: np.ndarray) -> float:
    """The reference elevation a run's datum setting asks for: None -> 0 (no shift,
    every input bit for bit as given); "auto" -> the whole metres below the lowest bed
    cell, floor(min(bed)), exactly representable; a number -> that number."""
    if datum is None:
        return 0.0
    if isinstance(datum, str):
        if datum == "auto":
            return float(np.floor(np.asarray(bed, dtype=np.float64).min()))
        raise ValueError(f"a datum is a number or 'auto', not {datum!r}")
    return float(datum)


def shift_bed(bed: np.ndarray, z_ref: float) -> np.ndarray:
    """The bed relative to z_ref, in float32; with z_ref = 0 the very same array."""
    if z_ref == 0.0:
        return bed
    return (np.asarray(bed, dtype=np.float32) - np.float32(z_ref)).astype(np.float32)


def unshift_bed(bed: np.ndarray, z_ref: float) -> np.ndarray:
    """A shifted bed back in true elevations; with z_ref = 0 the very same array."""
    if z_ref == 0.0:
        return bed
    return (np.asarray(bed, dtype=np.float32) + np.float32(z_ref)).astype(np.float32)


@dataclass(frozen=True)
class ShiftedElevations:
    """Everything absolute a run carries, in shifted coordinates."""

    z_ref: float
    bed: np.ndarray
    stage_curves: dict
    structures: list


def shift_for_datum(datum, bed: np.ndarray, stage_curves: dict | None = None,
                    structures: list | None = None) -> ShiftedElevations:
    """Resolve ``datum`` and move the bed, each fixed-stage curve's levels ((t, level)
    pairs per edge) and each structure's ``crest_m`` and ``target_stage_m`` by the
    same z_ref, in one place so that none of them can disagree with the bed."""
    z_ref = resolve_datum(datum, bed)
    curves = {edge: [(t, level - z_ref) for t, level in curve]
              for edge, curve in (stage_curves or {}).items()}
    moved = []
    for s in structures or []:
        target = None if s.target_stage_m is None else s.target_stage_m - z_ref
        moved.append(dataclasses.replace(s, crest_m=s.crest_m - z_ref, target_stage_m=target))
    return ShiftedElevations(z_ref=z_ref, bed=shift_bed(bed, z_ref), stage_curves=curves,
                             structures=moved)
