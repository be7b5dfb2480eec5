"""The datum shift's arithmetic (solver/core/test_datum.py:46-70, run.py:112-137); the
lake-at-rest experiment that justifies it runs on the GPU (tests/test_gates_gpu.py)."""

from __future__ import annotations

import numpy as np
import pytest

from rbs.core.datum import resolve_datum, shift_bed, shift_for_datum, unshift_bed
from rbs.processes.reservoir import Structure


def test_resolve_datum_modes():
    bed = np.array([[12.7, 15.2], [11.4, 19.0]], dtype=np.float32)
    assert resolve_datum(None, bed) == 0.0
    assert resolve_datum("auto", bed) == 11.0
    assert resolve_datum(250.0, bed) == 250.0
    with pytest.raises(ValueError, match="number or 'auto'"):
        resolve_datum("sea-level", bed)


def test_shift_round_trips_and_zero_is_the_same_array():
    yy, xx = np.mgrid[0:8, 0:8]
    bed = (500.0 + 0.05 * xx + 0.8 * np.sin(0.7 * xx) * np.cos(0.5 * yy)).astype(np.float32)
    shifted = shift_bed(bed, 500.0)
    assert float(np.abs(shifted).max()) < 2.0
    assert np.array_equal(unshift_bed(shifted, 500.0), bed)
    assert shift_bed(bed, 0.0) is bed and unshift_bed(bed, 0.0) is bed


def test_every_absolute_elevation_moves_with_the_bed():
    bed = np.full((10, 10), 312.4, np.float32)
    bed[5:] = 305.0
    dam = Structure(name="d", kind="dam", cells=[(4, j) for j in range(10)], crest_m=318.0,
                    release_rule="target_stage", target_stage_m=316.0, release_max_m3_s=5.0,
                    pool=(0, 0, 3, 9), outlet=(7, 2, 8, 5), interval_s=300.0)
    levee = Structure(name="l", kind="levee", cells=[(9, 0)], crest_m=306.5)
    el = shift_for_datum("auto", bed, {"south": [(0.0, 306.0), (3600.0, 307.5)]}, [dam, levee])
    assert el.z_ref == 305.0
    assert float(el.bed.min()) == 0.0 and float(el.bed.max()) == pytest.approx(7.4, abs=1e-5)
    assert el.stage_curves == {"south": [(0.0, 1.0), (3600.0, 2.5)]}
    assert el.structures[0].crest_m == 13.0 and el.structures[0].target_stage_m == 11.0
    assert el.structures[1].crest_m == 1.5 and el.structures[1].target_stage_m is None
    # the absolute levels over the shifted bed are the same depths as before
    assert el.stage_curves["south"][0][1] - float(el.bed[-1, 0]) == 306.0 - float(bed[-1, 0])
    none = shift_for_datum(None, bed, {}, [dam])
    assert none.z_ref == 0.0 and none.bed is bed and none.structures[0].crest_m == 318.0
