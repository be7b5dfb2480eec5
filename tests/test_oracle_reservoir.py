"""Reservoirs on the CPU: barriers and the release rules (host code) and the
oracle's restatement of the release, pinned by the reference's own tests
(solver/processes/test_reservoir.py) -- impoundment against a paired control,
overtopping, the mass-exact transfer, the pool cap, the closed-loop rule, the slow
clock through scheduled runs, and the deepest-cell gauge."""

from __future__ import annotations

import numpy as np
import pytest

from rbs.processes.reservoir import Structure, apply_barriers, check_boxes, gauge_cell
from rbs.scheduler import MultiRateScheduler, SlowProcess

NY, NX = 12, 8
DX = 10.0
DAM_ROW = 6
CREST = 5.0
POOL = (0, 0, 5, NX - 1)
OUTLET = (8, 2, 10, 5)


def _basin():
    z = np.zeros((NY, NX), np.float32)
    z[:DAM_ROW + 1] = 2.0
    z[DAM_ROW + 1:] = np.linspace(2.0, 0.0, NY - DAM_ROW - 1, dtype=np.float32)[:, None]
    return z


def _dam(**kw):
    base = dict(name="d", kind="dam", cells=[(DAM_ROW, j) for j in range(NX)], crest_m=CREST,
                release_rule="fixed", release_m3_s=1.0, pool=POOL, outlet=OUTLET, interval_s=300.0)
    base.update(kw)
    return Structure(**base)


_BARRIER_ONLY = dict(release_rule="none", pool=None, outlet=None)


def _rule(s: Structure) -> dict:
    return dict(release_rule=s.release_rule, release_m3_s=s.release_m3_s,
                target_stage_m=s.target_stage_m, crest_m=s.crest_m,
                release_max_m3_s=s.release_max_m3_s)


def _pool_depth(surface):
    bed = _basin()
    h = np.zeros((NY, NX), np.float32)
    h[:DAM_ROW] = np.maximum(surface - bed[:DAM_ROW], 0.0)
    return h


def _volume(h):
    return float(np.asarray(h, np.float64).sum()) * DX * DX


def _release(orc, st, dam, t, dt_slow):
    return orc.reservoir_release(st, _rule(dam), dam.pool, dam.outlet, gauge_cell(dam, st.z), t,
                                 dt_slow)


# --- 1. a barrier is geometry -----------------------------------------------------
def test_apply_barriers_raises_the_bed_and_never_lowers_it():
    bed = np.zeros((5, 5), np.float32)
    bed[2, 2] = 7.0
    out = apply_barriers(bed, [Structure(name="d", cells=[(2, 2), (2, 3)], crest_m=4.0)])
    assert out[2, 3] == 4.0 and out[2, 2] == 7.0 and bed[2, 3] == 0.0
    assert apply_barriers(bed, []) is bed
    with pytest.raises(ValueError, match="outside the 5x5 grid"):
        apply_barriers(bed, [Structure(name="d", cells=[(9, 0)], crest_m=1.0)])


def _run_pool(orc, bed, h0, steps=300):
    st = orc.State(bed, DX, depth=h0, manning=0.03)
    for _ in range(steps):
        orc.hllc_step(st, orc.hllc_compute_dt(st, 0.45, 5.0))
    return st.h


def test_barrier_impounds_and_the_same_water_escapes_without_it(orc):
    h0 = _pool_depth(4.5)
    kept = _run_pool(orc, apply_barriers(_basin(), [_dam(**_BARRIER_ONLY)]), h0)
    spilt = _run_pool(orc, _basin(), h0)
    v0 = _volume(h0)
    assert _volume(kept[:DAM_ROW]) > 0.9 * v0
    assert _volume(spilt[:DAM_ROW]) < 0.5 * _volume(kept[:DAM_ROW])
    assert np.isfinite(kept).all() and kept.min() >= 0.0


def test_a_pool_above_the_crest_overtops(orc):
    bed = apply_barriers(_basin(), [_dam(**_BARRIER_ONLY)])
    over = _run_pool(orc, bed, _pool_depth(CREST + 0.5))
    under = _run_pool(orc, bed, _pool_depth(CREST - 0.5))
    below_over, below_under = _volume(over[DAM_ROW + 1:]), _volume(under[DAM_ROW + 1:])
    assert below_over > 100.0
    assert below_over > 10.0 * max(below_under, 1e-9)


# --- 2. the transfer is mass-exact ----------------------------------------------------
def test_release_transfer_conserves_mass_exactly(orc):
    st = orc.State(apply_barriers(_basin(), [_dam()]), DX, depth=_pool_depth(4.5), manning=0.03)
    r0, c0, r1, c1 = POOL
    o0, p0, o1, p1 = OUTLET
    before = st.h.copy()
    _, stage, q, vol = _release(orc, st, _dam(), 300.0, 300.0)
    after = st.h
    banked = st.loss_volume(st.cell_area)
    assert q == 1.0 and stage == pytest.approx(4.5, abs=1e-6)
    assert vol == pytest.approx(300.0, rel=1e-6)
    pool_loss = _volume(before[r0:r1 + 1, c0:c1 + 1]) - _volume(after[r0:r1 + 1, c0:c1 + 1])
    assert pool_loss == pytest.approx(vol, rel=1e-9)
    out_gain = _volume(after[o0:o1 + 1, p0:p1 + 1]) - _volume(before[o0:o1 + 1, p0:p1 + 1])
    # the outlet gained what left the pool, less exactly the banked rounding
    assert out_gain + banked == pytest.approx(vol, rel=1e-12)
    assert (_volume(before) - _volume(after)) == pytest.approx(banked, abs=1e-9 * _volume(before))
    # each pool cell lost exactly f32(h * scale), scale = f32(V / stored)
    scale = np.float32(300.0 / _volume(before[r0:r1 + 1, c0:c1 + 1]))
    pb = before[r0:r1 + 1, c0:c1 + 1]
    assert np.array_equal(after[r0:r1 + 1, c0:c1 + 1], pb - pb * scale)


def test_release_is_capped_by_what_the_pool_holds(orc):
    h0 = _pool_depth(2.02)
    st = orc.State(apply_barriers(_basin(), [_dam()]), DX, depth=h0, manning=0.03)
    stored = _volume(h0)
    _, _, _, vol = _release(orc, st, _dam(release_m3_s=10_000.0), 300.0, 300.0)
    assert vol <= stored + 1e-9 and vol == pytest.approx(stored, rel=1e-6)
    assert float(st.h[:DAM_ROW + 1].min()) >= 0.0
    assert not st.h[:DAM_ROW].any()  # scale 1: every pool cell emptied exactly


def test_release_on_a_dry_pool_is_a_no_op(orc):
    st = orc.State(apply_barriers(_basin(), [_dam()]), DX, depth=0.0, manning=0.03)
    assert _release(orc, st, _dam(), 300.0, 300.0)[1:] == (None, 0.0, 0.0)


# --- 3. the rules behave ---------------------------------------------------------------
def test_target_stage_draws_the_pool_down_and_eases_off(orc):
    dam = _dam(release_rule="target_stage", target_stage_m=3.0, release_max_m3_s=8.0)
    st = orc.State(apply_barriers(_basin(), [dam]), DX, depth=_pool_depth(4.8), manning=0.03)
    recs = [_release(orc, st, dam, 300.0 * k, 300.0) for k in range(1, 25)]
    stages = [r[1] for r in recs]
    qs = [r[2] for r in recs]
    assert stages[0] > stages[-1] and qs[0] > qs[-1]
    assert max(qs) <= dam.release_max_m3_s + 1e-9
    assert stages[-1] >= dam.target_stage_m - 0.05 and stages[-1] < 4.8
    # the proportional rule in closed form at the first activation
    assert qs[0] == pytest.approx(8.0 * (stages[0] - 3.0) / (CREST - 3.0), rel=1e-12)


def test_structure_validation_is_the_references():
    with pytest.raises(ValueError, match="type must be one of"):
        _dam(kind="weir")
    with pytest.raises(ValueError, match="at least one barrier cell"):
        _dam(cells=[])
    with pytest.raises(ValueError, match="a levee is barrier geometry only"):
        _dam(kind="levee")
    with pytest.raises(ValueError, match="needs both a 'pool' box and an 'outlet'"):
        _dam(pool=None)
    with pytest.raises(ValueError, match="overlaps the pool"):
        _dam(outlet=(2, 2, 3, 3))
    with pytest.raises(ValueError, match="release_m3_s > 0"):
        _dam(release_m3_s=0.0)
    with pytest.raises(ValueError, match="below crest_m"):
        _dam(release_rule="target_stage", target_stage_m=6.0, release_max_m3_s=1.0)
    with pytest.raises(ValueError, match="interval_s must be > 0"):
        _dam(interval_s=0.0)
    with pytest.raises(ValueError, match="outside the"):
        check_boxes(_dam(pool=(0, 0, 5, NX + 40)), (NY, NX))
    assert _dam(outlet=(9, 3)).outlet == (9, 3, 9, 3)  # a single cell is a 1x1 box
    assert _dam().discharge_at(None) == 0.0


def test_the_gauge_is_the_deepest_cell_of_the_pool(orc):
    """test_reservoir.py:313-335: a film stranded high in the box is not a stage."""
    bed = _basin().copy()
    bed[0, 0] = 8.0
    bed[3, 3] = 1.0
    assert gauge_cell(_dam(), bed) == (3, 3)
    h = np.zeros((NY, NX), np.float32)
    h[0, 0] = 0.05
    st = orc.State(bed, DX, depth=h, manning=0.03)
    assert _release(orc, st, _dam(), 300.0, 300.0)[1] is None
    h[3, 3] = 1.5
    st = orc.State(bed, DX, depth=h, manning=0.03)
    assert _release(orc, st, _dam(), 300.0, 300.0)[1] == pytest.approx(2.5, abs=1e-5)


# --- 4. the slow clock is real -----------------------------------------------------------
def _scheduled_run(orc, dams, *, end=1200.0, dt_max=5.0, depth=1.0):
    """run.py:357-388 with the oracle: HLLC fast steps, each dam's rule a slow process."""
    bed = apply_barriers(_basin(), dams)
    st = orc.State(bed, DX, depth=depth, manning=0.035)
    led = orc.MassLedger.from_state(st)
    series = {}
    procs = []
    for d in dams:
        if d.release_rule == "none":
            continue
        series[d.name] = []
        g = gauge_cell(d, st.z)

        def adv(t, el, d=d, g=g):
            series[d.name].append(orc.reservoir_release(st, _rule(d), d.pool, d.outlet, g, t, el))

        procs.append(SlowProcess(name=d.name, interval=d.interval_s, advance=adv))
    sched = MultiRateScheduler(end_time=end, output_every=600.0, processes=procs)
    for tick in sched.ticks(lambda: orc.hllc_compute_dt(st, 0.45, dt_max)):
        orc.hllc_step(st, tick.dt, t=tick.t0)
        for proc, el in tick.due:
            proc.advance(tick.t1, el)
        if tick.is_output:
            led.record(st, tick.t1)
    return st, led, series


def test_release_lands_on_the_slow_clock_through_a_full_run(orc):
    st, led, series = _scheduled_run(orc, [_dam(interval_s=300.0, release_m3_s=2.0)])
    assert led.max_rel_error < orc.MASS_GATE
    times = [r[0] for r in series["d"]]
    assert times == pytest.approx([300.0, 600.0, 900.0, 1200.0])
    assert all(r[3] > 0.0 for r in series["d"])


def test_a_release_rule_actually_changes_the_run(orc):
    off, led_off, _ = _scheduled_run(orc, [_dam(**_BARRIER_ONLY)])
    on, led_on, _ = _scheduled_run(orc, [_dam(release_m3_s=4.0)])
    assert _volume(off.h[:DAM_ROW]) - _volume(on.h[:DAM_ROW]) > 0.0
    assert led_off.max_rel_error < orc.MASS_GATE and led_on.max_rel_error < orc.MASS_GATE


def test_the_slow_clock_is_independent_of_the_fast_timestep(orc):
    times = []
    for dt_max in (5.0, 1.0):
        _, _, series = _scheduled_run(orc, [_dam(interval_s=400.0, release_m3_s=2.0)],
                                      dt_max=dt_max)
        times.append([r[0] for r in series["d"]])
    assert times[0] == pytest.approx([400.0, 800.0, 1200.0])
    assert times[0] == pytest.approx(times[1])
