"""Point sources, hydrographs and stage curves in the oracle, pinned by the
reference's own known answers (processes/test_inflow.py) and closed forms."""

from __future__ import annotations

import numpy as np
import pytest


def _inject(orc, add: float, nsteps: int, depth: float = 1.0):
    """Kahan injection of a constant depth increment onto one cell of a 3x3 state."""
    st = orc.State(np.zeros((3, 3), np.float32), 100.0, depth=depth)
    comp = np.zeros(1, np.float32)
    adds = np.array([add], np.float32)
    for _ in range(nsteps):
        orc.inject_point_sources_c(st, [1], [1], adds, comp)
    return float(st.h[1, 1]), comp


def _plain(add: float, nsteps: int, depth: float = 1.0) -> float:
    h = np.float32(depth)
    for _ in range(nsteps):
        h = np.float32(h + np.float32(add))
    return float(h)


def test_sub_ulp_point_source_lands_only_when_compensated(orc):
    """test_inflow.py:143-158: 4e-8 m onto 1 m is under half an ulp, so the plain add
    lands exactly nothing in 2000 steps; compensated reaches the analytic total."""
    add, n = 4.0e-8, 2000
    assert _plain(add, n) == 1.0
    h, _ = _inject(orc, add, n)
    want = 1.0 + float(np.float32(add)) * n
    assert abs(h - want) <= 2 ** -23  # within one ulp of 1
    assert h > 1.0 + 7.5e-5


def test_flood_increment_error_ratio_exceeds_100(orc):
    """test_inflow.py:129-141: 0.027 m onto metre-deep water, 1500 steps; the
    compensated error is over 100x smaller than the plain float32 sum's."""
    add = np.float32(90.0 * 3.0 / 10000.0)
    n = 1500
    exact = 1.0 + float(add) * n
    plain = abs(_plain(float(add), n) - exact)
    h, comp = _inject(orc, float(add), n)
    comped = abs(h - exact)
    assert plain / max(comped, 1e-12) > 100.0
    assert np.any(comp != 0.0)  # the debt is live (fast-math canary, :161-174)


def test_constant_inflow_volume_known_answer(orc):
    """test_inflow.py:70-84: 3 m^3/s for 100 s in 10 s steps is 300 m^3."""
    hg = [(0.0, 3.0), (100.0, 3.0)]
    total = sum(orc.discharge_at(hg, t + 5.0) * 10.0 for t in np.arange(0.0, 100.0, 10.0))
    assert total == pytest.approx(300.0, rel=1e-12)


def test_discharge_is_piecewise_linear_and_zero_outside(orc):
    hg = [(0.0, 0.0), (200.0, 4.0), (600.0, 0.0)]
    assert orc.discharge_at(hg, -1.0) == 0.0
    assert orc.discharge_at(hg, 600.5) == 0.0
    assert orc.discharge_at(hg, 100.0) == pytest.approx(2.0)
    assert orc.discharge_at(hg, 400.0) == pytest.approx(2.0)
    assert orc.discharge_at(hg, 200.0) == 4.0
    # a step (repeated knot time): the first segment containing t wins at the knot
    step = [(0.0, 1.0), (10.0, 1.0), (10.0, 5.0), (20.0, 5.0)]
    assert orc.discharge_at(step, 10.0) == 1.0
    assert orc.discharge_at(step, 10.5) == 5.0
    # a single knot is a spike at its own time only
    assert orc.discharge_at([(5.0, 2.0)], 5.0) == 2.0
    assert orc.discharge_at([(5.0, 2.0)], 5.5) == 0.0
    # the midpoint rule integrates a linear ramp exactly
    vol = sum(orc.discharge_at(hg, t + 0.5) for t in range(0, 200))
    assert vol == pytest.approx(0.5 * 200.0 * 4.0, rel=1e-12)


def test_stage_is_held_outside_its_curve(orc):
    """boundaries.py:49-70: linear inside, the end values held outside."""
    curve = [(0.0, 0.5), (200.0, 0.9), (400.0, 0.4)]
    assert orc.stage_at(curve, -10.0) == 0.5
    assert orc.stage_at(curve, 1e6) == 0.4
    assert orc.stage_at(curve, 100.0) == pytest.approx(0.7)
    assert orc.stage_at(curve, 300.0) == pytest.approx(0.65)
    with pytest.raises(ValueError):
        orc.stage_at([], 0.0)
