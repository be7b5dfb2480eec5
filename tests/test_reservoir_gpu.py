"""Reservoir releases on the HIP path against the oracle's restatement, and the slow
clock in host-loop and device-driven runs (solver/processes/test_reservoir.py)."""

from __future__ import annotations

import numpy as np
import pytest

from test_oracle_reservoir import (CREST, DAM_ROW, DX, NX, NY, OUTLET, POOL, _basin, _dam,
                                   _pool_depth, _rule, _volume)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rbs():
    from rbs import _lib
    from rbs.core import hllc, local_inertial, massbalance, state
    from rbs.processes import reservoir

    _lib.load()
    return type("R", (), dict(lib=_lib, li=local_inertial, hllc=hllc, State=state.State,
                              mb=massbalance, res=reservoir))


def _pair(rbs, orc, dam, depth):
    bed = rbs.res.apply_barriers(_basin(), [dam])
    g = rbs.State.from_bed(bed, DX, depth=depth, manning=0.03)
    o = orc.State(bed, DX, depth=depth, manning=0.03)
    return g, o, rbs.res.ReservoirOperator(dam, g)


def _oracle_release(orc, o, dam, t, el):
    return orc.reservoir_release(o, _rule(dam), dam.pool, dam.outlet,
                                 _gauge(dam, o.z), t, el)


def _gauge(dam, bed):
    from rbs.processes.reservoir import gauge_cell

    return gauge_cell(dam, bed)


def test_a_release_matches_the_oracle_and_closes_the_ledger(rbs, orc):
    dam = _dam()
    g, o, op = _pair(rbs, orc, dam, _pool_depth(4.5))
    led = rbs.mb.MassLedger.from_state(g)
    before = g.h.numpy()
    rec = op.advance(300.0, 300.0)
    want = _oracle_release(orc, o, dam, 300.0, 300.0)
    after = g.h.numpy()
    assert np.array_equal(after, o.h)
    assert (rec.time, rec.stage, rec.discharge_m3_s) == (want[0], want[1], want[2])
    assert rec.volume_m3 == pytest.approx(want[3], rel=1e-12)
    assert rec.volume_m3 == pytest.approx(300.0, rel=1e-6)
    banked = g.loss_volume()
    assert banked == pytest.approx(o.loss_volume(o.cell_area), rel=1e-9, abs=1e-15)
    assert (_volume(before) - _volume(after)) == pytest.approx(banked, abs=1e-9 * _volume(before))
    assert led.record(g, 300.0).rel_error < 1e-12


def test_capped_dry_and_closed_loop_releases_match_the_oracle(rbs, orc):
    g, o, op = _pair(rbs, orc, _dam(release_m3_s=10_000.0), _pool_depth(2.02))
    rec = op.advance(300.0, 300.0)
    want = _oracle_release(orc, o, _dam(release_m3_s=10_000.0), 300.0, 300.0)
    assert np.array_equal(g.h.numpy(), o.h) and rec.volume_m3 == pytest.approx(want[3], rel=1e-12)
    assert not g.h.numpy()[:DAM_ROW].any()

    g, o, op = _pair(rbs, orc, _dam(), 0.0)
    rec = op.advance(300.0, 300.0)
    assert rec.stage is None and rec.discharge_m3_s == 0.0 and rec.volume_m3 == 0.0

    dam = _dam(release_rule="target_stage", target_stage_m=3.0, release_max_m3_s=8.0)
    g, o, op = _pair(rbs, orc, dam, _pool_depth(4.8))
    for k in range(1, 25):
        rec = op.advance(300.0 * k, 300.0)
        want = _oracle_release(orc, o, dam, 300.0 * k, 300.0)
        assert rec.stage == want[1] and rec.discharge_m3_s == want[2]
        assert np.array_equal(g.h.numpy(), o.h), k
    assert op.records[0].stage > op.records[-1].stage >= 3.0 - 0.05


def test_the_gauge_is_the_deepest_cell(rbs):
    bed = _basin().copy()
    bed[0, 0] = 8.0
    bed[3, 3] = 1.0
    h = np.zeros((NY, NX), np.float32)
    h[0, 0] = 0.05
    st = rbs.State.from_bed(bed, DX, depth=h, manning=0.03)
    op = rbs.res.ReservoirOperator(_dam(), st)
    assert op._gauge == (3, 3) and op.pool_stage() is None
    assert op.advance(300.0, 300.0).stage is None
    h[3, 3] = 1.5
    st.h.assign(h)
    assert op.pool_stage() == pytest.approx(2.5, abs=1e-5)
    with pytest.raises(ValueError, match="outside the"):
        rbs.res.ReservoirOperator(_dam(pool=(0, 0, 5, NX + 40)), st)
    assert [o.structure.name for o in rbs.res.build_operators(
        [_dam(release_rule="none", pool=None, outlet=None), _dam(name="live")], st)] == ["live"]


def test_hllc_run_with_a_release_rule_matches_the_oracle(rbs, orc):
    """test_reservoir.py:266-283 through the HIP path: releases on the slow clock,
    the mass gate, and the oracle's scheduled run."""
    from rbs.scheduler import MultiRateScheduler
    from test_oracle_reservoir import _scheduled_run

    dam = _dam(interval_s=300.0, release_m3_s=2.0)
    o, o_led, o_series = _scheduled_run(orc, [dam])
    bed = rbs.res.apply_barriers(_basin(), [dam])
    g = rbs.State.from_bed(bed, DX, depth=1.0, manning=0.035)
    op = rbs.res.ReservoirOperator(dam, g)
    led = rbs.mb.MassLedger.from_state(g)
    sched = MultiRateScheduler(end_time=1200.0, output_every=600.0,
                               processes=[op.as_slow_process()])
    for tick in sched.ticks(lambda: rbs.hllc.compute_dt(g, 0.45, 5.0)):
        rbs.hllc.step(g, tick.dt, t=tick.t0)
        for proc, el in tick.due:
            proc.advance(tick.t1, el)
        if tick.is_output:
            led.record(g, tick.t1)
    assert [r.time for r in op.records] == pytest.approx([300.0, 600.0, 900.0, 1200.0])
    assert [r.time for r in op.records] == [r[0] for r in o_series["d"]]
    for r, w in zip(op.records, o_series["d"]):
        assert r.volume_m3 == pytest.approx(w[3], rel=1e-5)
    assert np.abs(g.h.numpy() - o.h).max() <= 1e-4
    assert led.max_rel_error < 1e-6
    assert float(bed[DAM_ROW, 0]) == CREST


def test_device_driven_run_with_a_reservoir_equals_the_host_loop(rbs):
    """A device-driven run stops at each activation for the release; the result is the
    host tick loop's, bit for bit, and so is the release series."""
    from rbs.driver import DeviceRun
    from rbs.scheduler import MultiRateScheduler

    dam = _dam(release_rule="target_stage", target_stage_m=3.5, release_max_m3_s=6.0,
               interval_s=120.0)
    bed = rbs.res.apply_barriers(_basin(), [dam])
    runs = []
    for driven in (False, True):
        st = rbs.State.from_bed(bed, DX, depth=_pool_depth(4.6), manning=0.035)
        st.set_open_boundaries({"south": "open"})
        op = rbs.res.ReservoirOperator(dam, st)
        sched = MultiRateScheduler(end_time=1200.0, output_every=300.0,
                                   processes=[op.as_slow_process()])
        if driven:
            run = DeviceRun(st, "local_inertial", sched, alpha=0.7, dt_max=5.0)
            for seg in sched.segments():
                run.run_until(seg.t_stop)
                for proc, el in seg.due:
                    proc.advance(seg.t_stop, el)
        else:
            for tick in sched.ticks(lambda: rbs.li.compute_dt(st, 0.7, 5.0)):
                rbs.li.step(st, tick.dt, t=tick.t0)
                for proc, el in tick.due:
                    proc.advance(tick.t1, el)
        runs.append((st.h.numpy(), op.series, st.loss_volume()))
    (h0, s0, l0), (h1, s1, l1) = runs
    assert len(s0) == 10 and all(r["volume_m3"] > 0 for r in s0)
    assert s0 == s1
    assert np.array_equal(h0, h1)
    assert l0 == l1 and float(h0[OUTLET[0]:OUTLET[2] + 1].max()) > 0.0
    assert POOL[2] < DAM_ROW
