"""Bedload integral, Exner activation and point sources on the HIP path, against
the oracle (and the reference's sediment ledger gate)."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rbs():
    from rbs import _lib
    from rbs.core import local_inertial, massbalance, sediment, state
    from rbs.processes import inflow, morphology

    _lib.load()
    return type("R", (), dict(li=local_inertial, State=state.State, sed=sediment,
                              mb=massbalance, inflow=inflow, morph=morphology))


def _reach(ny=24, nx=160, dx=5.0):
    x = np.arange(nx)[None, :] * np.ones((ny, 1))
    rng = np.random.default_rng(3)
    return (3e-3 * dx * (nx - x) + 0.05 * rng.uniform(0, 1, (ny, nx))).astype(np.float32)


def test_transport_integral_and_exner_match_the_oracle(rbs, orc):
    z = _reach()
    d50 = np.random.default_rng(1).uniform(5e-4, 4e-3, z.shape).astype(np.float32)
    g = rbs.State.from_bed(z, 5.0, depth=0.5, manning=0.03)
    o = orc.State(z, 5.0, depth=0.5, manning=0.03)
    rbs.sed.arm_sediment(g, d50, 0.4)
    o.arm_sediment(d50)
    for _ in range(150):
        dt = rbs.li.compute_dt(g, 0.7, 2.0)
        assert abs(dt - orc.li_compute_dt(o, 0.7, 2.0)) <= 1e-6 * dt
        rbs.li.step(g, dt)
        orc.li_step(o, dt)
    sed = g.sediment
    np.testing.assert_array_equal(sed.qs_int_x.numpy(), o.qs_int_x)  # one power sequence: bitwise
    np.testing.assert_array_equal(sed.qs_int_y.numpy(), o.qs_int_y)
    assert np.abs(g.h.numpy() - o.h).max() <= 1e-5
    # one activation on the device vs the oracle's exner on the device's integrals
    qx, qy = sed.qs_int_x.numpy(), sed.qs_int_y.numpy()
    comp_before = sed.qs_comp_x.numpy()
    lo = np.full(z.shape, -np.inf, np.float32)
    hi = np.full(z.shape, np.inf, np.float32)
    lo[:, :3] = 0.0
    hi[:, :3] = 0.0  # frozen inlet: banks what it refuses
    proc = rbs.morph.MorphologyProcess(g, 300.0, dz_lo=lo, dz_hi=hi)
    led = rbs.mb.SedimentLedger.from_state(g)
    proc.advance(300.0, 300.0)
    dz = np.zeros(z.shape, np.float64)
    un = np.zeros(z.shape, np.float64)
    orc.exner_update(qx, qy, lo, hi, 5.0, 0.4, dz, un)
    zo = np.zeros_like(z)
    orc.rebuild_bed(z, dz, zo)
    # the same float64 arithmetic in numpy, as a second witness
    want = -np.float64(np.float32(1.0 / 0.6)) * ((qx[:, 1:].astype(np.float64) - qx[:, :-1])
                                                 + (qy[1:, :].astype(np.float64) - qy[:-1, :])) \
        / np.float64(np.float32(5.0))
    np.testing.assert_array_equal(dz, np.clip(want, lo, hi))
    np.testing.assert_array_equal(sed.dz_cum.numpy(), dz)
    np.testing.assert_array_equal(sed.dz_unapplied.numpy(), un)
    np.testing.assert_array_equal(g.z.numpy(), zo)
    assert np.all(sed.qs_int_x.numpy() == 0.0)  # integral cleared ...
    assert np.array_equal(sed.qs_comp_x.numpy(), comp_before)  # ... its debt kept
    rec = led.record(g, 300.0)
    assert rec.rel_error < 1e-11  # massbalance.py:178 SEDIMENT_GATE
    assert rec.gross_volume > 0
    # the device bed-change totals against numpy on the downloaded fields
    tot = sed.totals()
    assert tot["min"] == dz.min() and tot["max"] == dz.max()
    assert tot["sum"] == pytest.approx(dz.sum(), rel=1e-12, abs=1e-18)
    assert tot["sum_abs"] == pytest.approx(np.abs(dz).sum(), rel=1e-12)
    assert tot["unapplied"] == pytest.approx(un.sum(), rel=1e-12, abs=1e-18)
    assert tot == sed.totals()  # fixed-order reduction: the same bits again


def test_point_sources_match_the_oracle_and_the_ledger(rbs, orc):
    z = np.zeros((20, 30), np.float32)
    g = rbs.State.from_bed(z, 10.0, depth=0.2, manning=0.03)
    o = orc.State(z, 10.0, depth=0.2, manning=0.03)
    srcs = [rbs.inflow.Inflow((3, 4), [(0.0, 0.0), (50.0, 5.0), (120.0, 1.0)]),
            rbs.inflow.Inflow((15, 25), [(10.0, 2.0), (100.0, 2.0)])]
    inj = rbs.inflow.InflowInjector(srcs, g)
    led = rbs.mb.MassLedger.from_state(g)
    comp = np.zeros(2, np.float32)
    rows, cols = [3, 15], [4, 25]
    t = 0.0
    for _ in range(60):
        dt = rbs.li.compute_dt(g, 0.7, 3.0)
        assert abs(dt - orc.li_compute_dt(o, 0.7, 3.0)) <= 1e-6 * dt
        vol = inj.apply(g, t, dt)
        add = np.array([orc.discharge_at(s.hydrograph, t + 0.5 * dt) * dt / 100.0 for s in srcs])
        add32 = add.astype(np.float32)
        assert vol == pytest.approx(float(add32.astype(np.float64).sum()) * 100.0, rel=1e-15)
        orc.inject_point_sources_c(o, rows, cols, add32, comp)
        led.add_inflow(vol)
        rbs.li.step(g, dt, t=t)
        orc.li_step(o, dt)
        t += dt
    assert np.abs(g.h.numpy() - o.h).max() <= 1e-6
    assert led.record(g, t).rel_error < 1e-6


def test_device_driven_point_sources_match_host_loop(rbs):
    from rbs.driver import DeviceRun
    from rbs.scheduler import MultiRateScheduler

    z = _reach(16, 64)
    hg = [(0.0, 0.5), (60.0, 4.0), (180.0, 0.5)]
    sched = MultiRateScheduler(end_time=240.0, output_every=60.0, events=[0.0, 60.0, 180.0])
    host = rbs.State.from_bed(z, 5.0, depth=0.05, manning=0.03)
    dev = rbs.State.from_bed(z, 5.0, depth=0.05, manning=0.03)
    inj_h = rbs.inflow.InflowInjector([rbs.inflow.Inflow((8, 60), hg)], host)
    rbs.inflow.InflowInjector([rbs.inflow.Inflow((8, 60), hg)], dev)
    led_h = rbs.mb.MassLedger.from_state(host)
    led_d = rbs.mb.MassLedger.from_state(dev)
    n_h = 0
    for tick in sched.ticks(lambda: rbs.li.compute_dt(host, 0.7, 10.0)):
        led_h.add_inflow(inj_h.apply(host, tick.t0, tick.dt))
        rbs.li.step(host, tick.dt, t=tick.t0)
        n_h += 1
    run = DeviceRun(dev, "local_inertial", sched, alpha=0.7, dt_max=10.0)
    n_d = sum(run.run_until(s.t_stop) for s in sched.segments())
    assert n_d == n_h
    assert np.array_equal(host.h.numpy(), dev.h.numpy())
    rh, rd = led_h.record(host, 240.0), led_d.record(dev, 240.0)
    assert rd.inflow_cum == pytest.approx(rh.inflow_cum, rel=1e-13)
    assert rd.rel_error < 1e-6
