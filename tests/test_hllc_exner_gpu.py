"""HLLC + Exner on the HIP path: the stage kernels' bedload integral against the
oracle, and the bed-wave and threshold invariants run on the device."""

from __future__ import annotations

import math

import numpy as np
import pytest

from test_oracle_hllc_exner import EPS_T, Reach, drive, xcorr_lag

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rbs():
    from rbs import _lib
    from rbs.core import hllc, massbalance, sediment, state
    from rbs.processes import inflow, morphology

    _lib.load()
    return type("R", (), dict(lib=_lib, hllc=hllc, State=state.State, sed=sediment,
                              mb=massbalance, inflow=inflow, morph=morphology))


def test_stage_transport_integral_matches_the_oracle(rbs, orc):
    rng = np.random.default_rng(8)
    ny, nx, dx = 40, 96, 5.0
    yy, xx = np.mgrid[0:ny, 0:nx]
    z = (0.01 * dx * (nx - xx) + 0.2 * rng.uniform(0, 1, (ny, nx))).astype(np.float32)
    h = np.where(xx < nx // 2, 1.2, 0.3).astype(np.float32)
    d50 = rng.uniform(5e-4, 3e-3, (ny, nx)).astype(np.float32)
    g = rbs.State.from_bed(z, dx, depth=h, manning=0.03)
    o = orc.State(z, dx, depth=h, manning=0.03)
    for s in (g, o):
        s.set_open_boundaries({"east": "open", "west": "closed"})
    rbs.sed.arm_sediment(g, d50, 0.4)
    o.arm_sediment(d50)
    t = 0.0
    for _ in range(150):
        dt = rbs.hllc.compute_dt(g, 0.45, 1.0)
        assert abs(dt - orc.hllc_compute_dt(o, 0.45, 1.0)) <= 1e-6 * dt
        rbs.hllc.step(g, dt, t=t)
        orc.hllc_step(o, dt, t=t)
        t += dt
    assert np.abs(g.h.numpy() - o.h).max() <= 1e-5
    sed = g.sediment
    np.testing.assert_array_equal(sed.qs_int_x.numpy(), o.qs_int_x)  # one power sequence: bitwise
    np.testing.assert_array_equal(sed.qs_int_y.numpy(), o.qs_int_y)
    assert np.abs(o.qs_int_x).max() > 0.0


def _drive_device(rbs, fx: Reach, n_act: int):
    """The fixture driver of test_oracle_hllc_exner on the device."""
    st = rbs.State.from_bed(fx.bed(), fx.dx, depth=fx.depth, manning=fx.manning)
    st.set_open_boundaries({"east": "fixed_stage"}, {"east": [(0.0, fx.outlet_bed + fx.depth)]})
    inj = rbs.inflow.InflowInjector(
        [rbs.inflow.Inflow((0, 0), [(0.0, fx.q * fx.dx), (1.0e9, fx.q * fx.dx)])], st)
    led = rbs.mb.MassLedger.from_state(st)
    lo = np.full((1, fx.nx), -np.inf, np.float32)
    hi = np.full((1, fx.nx), np.inf, np.float32)
    lo[0, [0, -1]] = hi[0, [0, -1]] = 0.0
    morph = None
    end = fx.warmup_s + n_act * fx.interval_s
    t, acts = 0.0, 0
    from rbs.scheduler import fill_span

    while t < end - EPS_T:
        dt = rbs.hllc.compute_dt(st, 0.45, 5.0)
        edge = fx.warmup_s if t < fx.warmup_s - EPS_T else fx.warmup_s + (acts + 1) * fx.interval_s
        dt = fill_span(min(edge, end) - t, dt)
        if morph is None and t >= fx.warmup_s - EPS_T:
            rbs.sed.arm_sediment(st, fx.d50, fx.porosity)
            morph = rbs.morph.MorphologyProcess(st, fx.interval_s, dz_lo=lo, dz_hi=hi)
        led.add_inflow(inj.apply(st, t, dt))
        rbs.hllc.step(st, dt, t=t)
        t += dt
        if morph is not None and t >= fx.warmup_s + (acts + 1) * fx.interval_s - EPS_T:
            morph.advance(t, fx.interval_s)
            acts += 1
    led.record(st, t)
    return st, led, acts


def test_the_bump_migrates_at_the_swe_celerity_on_the_device(rbs, orc):
    fx = Reach()
    st, led, acts = _drive_device(rbs, fx, fx.activations)
    o, o_led, o_dz, _, _ = drive(orc, fx, n_act=fx.activations)
    h = st.h.numpy()
    interior = slice(30, fx.nx - 30)
    hh, q = float(np.median(h[0, interior])), float(np.median(st.hu.numpy()[0, interior]))
    froude = q / hh / math.sqrt(9.81 * hh)
    c_swe = fx.celerity_at(q, hh) / (1.0 - froude ** 2)
    z = st.z.numpy()
    lo, hi = fx.window
    final = np.asarray(z[0], np.float64) - fx.plane()
    lag = xcorr_lag(fx.bump(), final, lo, hi, int(math.ceil(2 * fx.migration_cells)))
    celerity = lag * fx.dx / (acts * fx.interval_s)
    print(f"\n[device hllc bed wave] {celerity / c_swe:.3f} of c_b/(1-Fr^2); "
          f"max |z - oracle z| {np.abs(z - o.z).max():.2e} m; mass {led.max_rel_error:.1e}")
    assert acts == fx.activations
    assert celerity == pytest.approx(c_swe, rel=0.20)
    assert np.abs(z - o.z).max() < 2e-4  # the oracle's bed, to a small part of the bump
    assert led.max_rel_error < 1e-5
    dz = st.sediment.dz_cum.numpy()
    un = st.sediment.dz_unapplied.numpy()
    assert abs(float(dz.sum() + un.sum())) / float(np.abs(dz).sum()) < 1e-11


def test_under_the_threshold_the_device_moves_no_bed(rbs):
    fx = Reach().at_shields(0.90)
    st, led, acts = _drive_device(rbs, fx, 20)
    assert acts == 20
    assert not st.sediment.dz_cum.numpy().any()
    assert np.array_equal(st.z.numpy(), fx.bed())


def test_a_lake_at_rest_with_sediment_armed_is_the_unarmed_lake(rbs):
    rng = np.random.default_rng(3)
    z = (2.0 * rng.uniform(0, 1, (24, 40))).astype(np.float32)
    h = np.maximum(np.float32(3.0) - z, 0).astype(np.float32)
    plain = rbs.State.from_bed(z, 5.0, depth=h, manning=0.03)
    armed = rbs.State.from_bed(z, 5.0, depth=h, manning=0.03)
    rbs.sed.arm_sediment(armed, np.full(z.shape, 1e-4, np.float32), 0.4)
    for _ in range(200):
        dt = rbs.hllc.compute_dt(plain, 0.45, 1.0)
        assert dt == rbs.hllc.compute_dt(armed, 0.45, 1.0)
        rbs.hllc.step(plain, dt)
        rbs.hllc.step(armed, dt)
    assert np.array_equal(plain.h.numpy(), armed.h.numpy())
    assert np.array_equal(plain.hu.numpy(), armed.hu.numpy())
    assert not armed.sediment.qs_int_x.numpy().any()
