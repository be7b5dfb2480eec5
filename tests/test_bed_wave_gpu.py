"""The reference's local-inertial morphology gates on the HIP path:
validation/test_bed_wave.py (the bed-wave celerity and its fences),
validation/test_morphology_gates.py:85-155 (the Shields threshold pair) and
validation/test_clamp_ripple.py (equal steps to every sync point: a frequent sync
cadence must not ripple a uniform reach), the last both through the host loop and
through device-driven runs, whose fill_span runs on the GPU.

The fixture is tests/flume.py.  The gate widths are the reference's.
"""

from __future__ import annotations

import math
from dataclasses import replace

import numpy as np
import pytest

from flume import DT_MAX, Flume, drive, measure
from rbs.core.sediment import SHIELDS_CRITICAL, morphological_courant, shields_from_flow

MASS_GATE = 1e-6  # massbalance.py:36


@pytest.fixture(scope="module")
def rbs():
    from rbs import _lib
    from rbs.core import local_inertial, massbalance, sediment, state
    from rbs.processes import inflow, morphology

    _lib.load()
    return type("R", (), dict(lib=_lib, li=local_inertial, State=state.State, sed=sediment,
                              mb=massbalance, inflow=inflow, morph=morphology))


def test_the_flume_is_sized_for_the_law():
    """test_bed_wave.py:76-110, arithmetic only (no GPU): every sizing constraint."""
    fx = Flume()
    lo, hi = fx.window
    sigma_cells = fx.sigma_m / fx.dx
    assert fx.shields / SHIELDS_CRITICAL > 3.0 and fx.shields < 1.0
    assert fx.sigma_m / (fx.depth / fx.slope) < 0.05
    assert fx.froude < 0.6
    assert sigma_cells >= 5.0
    assert fx.amp < 0.02 * fx.depth
    assert fx.courant < 0.25
    assert lo > fx.pinned + sigma_cells and hi < fx.nx - fx.pinned
    assert fx.bump_cell + fx.migration_cells + 2 * sigma_cells < hi
    assert fx.migration_cells * fx.dx > 2.0 * fx.sigma_m
    assert fx.celerity < 0.01 * math.sqrt(9.81 * fx.depth)
    assert morphological_courant(fx.celerity, 900.0, fx.dx) > 1.0
    # the design point's printed figures (bedwave.py:25-43)
    assert fx.depth == pytest.approx(1.496, abs=5e-4)
    assert fx.froude == pytest.approx(0.44, abs=5e-3)
    assert fx.celerity == pytest.approx(8.63e-3, rel=2e-3)


@pytest.mark.gpu
def test_the_design_point_is_what_the_solver_delivers(rbs):
    """test_bed_wave.py:113-156: morphology off, a few intervals past the warm-up."""
    fx = Flume()
    res = drive(rbs, fx, morphology=False, end_s=fx.warmup_s + 3.0 * fx.interval_s)
    h, q = res.median_depth(fx.interior), res.median_q(fx.interior)
    theta = float(shields_from_flow(q, h, fx.manning, fx.d50))
    c_flow = fx.celerity_at(q, h)
    print(f"\n[flume design point] h {h:.4f} m (design {fx.depth:.4f}), q {q:.4f}, theta "
          f"{theta:.4f} ({fx.shields:.4f}), c_b {c_flow:.4e} ({fx.celerity:.4e}), "
          f"mass {res.mass_rel_error:.2e}")
    assert res.mass_rel_error < MASS_GATE
    assert h == pytest.approx(fx.depth, rel=0.01)
    assert q == pytest.approx(fx.q, rel=0.01)
    assert theta == pytest.approx(fx.shields, rel=0.02)
    assert c_flow == pytest.approx(fx.celerity, rel=0.05)
    assert (res.bed == fx.bed()[0]).all() and (res.dz_cum == 0.0).all()


@pytest.mark.gpu
def test_the_bump_migrates_at_the_bed_wave_celerity(rbs):
    """test_bed_wave.py:159-211: the celerity gate, +-20 % on the shape estimator
    against c_b at the achieved flow; the local-inertial scheme drops advection, so this
    is the rigid-lid celerity (HLLC's is c_b / (1 - Fr^2), test_hllc_exner_gpu.py)."""
    fx = Flume()
    res = drive(rbs, fx)
    mig = measure(fx, res.bed, res.activations * fx.interval_s)
    c_flow = fx.celerity_at(res.median_q(fx.interior), res.median_depth(fx.interior))
    xc, pk, ce = (mig.speed(c) / c_flow for c in (mig.xcorr_cells, mig.peak_cells,
                                                    mig.centroid_cells))
    print(f"\n[LI bed wave] {res.steps} steps, {res.activations} activations: xcorr {xc:.3f} "
          f"c_b, crest {pk:.3f}, centroid {ce:.3f}; crest kept {mig.retained:.2f}, "
          f"signal/background {mig.signal_to_background:.1f}; peak Courant "
          f"{res.courant:.3f} (design {fx.courant:.3f}); mass {res.mass_rel_error:.2e}")
    assert res.mass_rel_error < MASS_GATE
    assert res.activations == fx.activations
    assert 0.5 * fx.courant < res.courant < 2.0 * fx.courant and res.courant < 0.25
    assert 0.5 < mig.retained < 1.05
    assert mig.signal_to_background > 5.0
    assert xc == pytest.approx(1.0, abs=0.20)
    assert 0.6 < pk < 1.4 and 0.5 < ce < 1.4


@pytest.mark.gpu
def test_the_final_bed_does_not_depend_on_the_interval(rbs):
    """test_bed_wave.py:222-270: 45 s and 90 s over the same 4680 s of morphology."""
    coarse, fine = replace(Flume(), interval_s=90.0), Flume()
    end = fine.warmup_s + 4680.0
    rc, rf = drive(rbs, coarse, end_s=end), drive(rbs, fine, end_s=end)
    dep_c, dep_f = rc.bed - coarse.plane(), rf.bed - fine.plane()
    diff = float(np.abs(dep_f - dep_c).max())
    signal = float(np.abs(dep_f).max())
    print(f"\n[interval independence] 90 s: {rc.activations} activations, Cr {rc.courant:.3f}; "
          f"45 s: {rf.activations}, Cr {rf.courant:.3f}; max |bed difference| "
          f"{1000 * diff:.4f} mm against {1000 * signal:.4f} mm")
    assert rc.activations * 90.0 == pytest.approx(rf.activations * 45.0)
    assert rc.mass_rel_error < MASS_GATE and rf.mass_rel_error < MASS_GATE
    assert rc.courant < 1.0 and rf.courant < 1.0
    assert diff < 0.05 * signal


@pytest.mark.gpu
def test_a_flat_reach_stays_flat(rbs):
    """test_bed_wave.py:273-306: no bump, uniform transport, no divergence."""
    fx = replace(Flume(), amp=0.0)
    res = drive(rbs, fx)
    spurious = float(np.abs(res.dz_cum).max())
    print(f"\n[flat reach] {res.activations} activations: spurious bed change "
          f"{1000 * spurious:.4f} mm, depth {res.median_depth(fx.interior):.4f} m")
    assert res.mass_rel_error < MASS_GATE
    assert res.median_depth(fx.interior) == pytest.approx(fx.depth, rel=0.01)
    assert spurious < 0.05 * Flume().amp


@pytest.mark.gpu
def test_an_over_courant_interval_is_caught(rbs):
    """test_bed_wave.py:309-352: at 900 s the morphological Courant number the process
    measures exceeds 1, and the bump grows (the artefact's signature)."""
    fx = replace(Flume(), interval_s=900.0)
    res = drive(rbs, fx)
    mig = measure(fx, res.bed, res.activations * fx.interval_s)
    print(f"\n[over-Courant] {res.activations} activations at 900 s: peak Courant "
          f"{res.courant:.3f}, crest kept {mig.retained:.3f}, mass {res.mass_rel_error:.2e}")
    assert res.mass_rel_error < MASS_GATE
    assert res.courant > 1.0 and Flume().courant < 1.0
    assert mig.retained > 1.0


@pytest.mark.gpu
def test_free_ends_grow_sills_pinned_ends_do_not(rbs):
    """test_bed_wave.py:355-396: the flume's sediment boundary condition."""
    short = replace(Flume(), migration_cells=4.0)
    free = drive(rbs, replace(short, pinned=0))
    pinned = drive(rbs, short)
    print(f"\n[ends] free: inlet {free.dz_cum[0]:+.4f} m outlet {free.dz_cum[-1]:+.4f} m depth "
          f"{free.median_depth(short.interior):.4f}; pinned: depth "
          f"{pinned.median_depth(short.interior):.4f}, banked {pinned.banked:+.4f} m")
    assert free.mass_rel_error < MASS_GATE and pinned.mass_rel_error < MASS_GATE
    assert free.dz_cum[0] < -0.5 and free.dz_cum[-1] > 0.5
    assert free.median_depth(short.interior) > 1.1 * short.depth
    assert free.banked == 0.0
    assert pinned.dz_cum[0] == 0.0 and pinned.dz_cum[-1] == 0.0
    assert pinned.median_depth(short.interior) == pytest.approx(short.depth, rel=0.01)
    assert pinned.banked > 0.0


def _threshold(rbs, ratio: float):
    fx = Flume().at_shields(ratio)
    res = drive(rbs, fx, end_s=fx.warmup_s + 20 * fx.interval_s)
    theta = np.asarray(shields_from_flow(res.face_q[1:-1][:-1], res.depth[1:-1], fx.manning,
                                         fx.d50))
    return fx, res, theta


@pytest.mark.gpu
def test_just_under_the_threshold_no_bed_moves(rbs):
    """test_morphology_gates.py:85-117: at 0.9 theta_c, crest included, the bed change
    is bit-exact zero after 20 activations."""
    fx, res, theta = _threshold(rbs, 0.90)
    print(f"\n[under threshold] theta {theta.min() / SHIELDS_CRITICAL:.3f}.."
          f"{theta.max() / SHIELDS_CRITICAL:.3f} theta_c; "
          f"{int((res.dz_cum != 0).sum())} cells moved")
    assert res.mass_rel_error < MASS_GATE and res.activations == 20
    assert theta.max() < SHIELDS_CRITICAL
    assert (res.dz_cum == 0.0).all()


@pytest.mark.gpu
def test_just_over_the_threshold_the_bed_moves(rbs):
    """test_morphology_gates.py:120-155: at 1.2 theta_c the whole reach transports."""
    fx, res, theta = _threshold(rbs, 1.20)
    moved = int((res.dz_cum != 0.0).sum())
    print(f"\n[over threshold] theta {theta.min() / SHIELDS_CRITICAL:.3f}.."
          f"{theta.max() / SHIELDS_CRITICAL:.3f} theta_c; {moved} of {fx.nx} cells moved")
    assert res.mass_rel_error < MASS_GATE
    assert theta.min() > SHIELDS_CRITICAL
    assert moved >= fx.nx - 4
    assert res.dz_cum[0] == 0.0 and res.dz_cum[-1] == 0.0
    assert np.abs(res.dz_cum).max() > 1e-3


# --- the clamp ripple (test_clamp_ripple.py) --------------------------------------------
CURVATURE_MM = 5.0
DT_JUMP = 0.01
END_S = 2400.0
CADENCES = (45.0, 11.25)


def _uniform_reach(rbs):
    fx = replace(Flume(), amp=0.0)
    st = rbs.State.from_bed(fx.bed(), fx.dx, depth=fx.depth, manning=fx.manning)
    st.set_open_boundaries({"east": "open"})
    inj = rbs.inflow.InflowInjector(
        [rbs.inflow.Inflow((0, 0), [(0.0, fx.q * fx.dx), (1.0e9, fx.q * fx.dx)])], st)
    return fx, st, inj


def _host_scheduled(rbs, cadence):
    """The uniform reach under the scheduler's ticks at a sync cadence (None: only the
    end), host loop (test_clamp_ripple.py:188-227)."""
    from rbs.scheduler import MultiRateScheduler

    fx, st, inj = _uniform_reach(rbs)
    led = rbs.mb.MassLedger.from_state(st)
    sched = MultiRateScheduler(end_time=END_S, output_every=END_S if cadence is None else cadence)
    dts, raws = [], []

    def dt_fn():
        raws.append(rbs.li.compute_dt(st, 0.7, DT_MAX))
        return raws[-1]

    worst = 0.0
    for tick in sched.ticks(dt_fn):
        led.add_inflow(inj.apply(st, tick.t0, tick.dt))
        rbs.li.step(st, tick.dt, t=tick.t0)
        dts.append(tick.dt)
        worst = max(worst, tick.dt / raws[-1])
    led.record(st, END_S)
    h = st.h.numpy()[0].astype(np.float64)
    st.close()
    return fx, h, np.asarray(dts), led.max_rel_error, worst


def _curvature_mm(fx, h) -> float:
    return 1000.0 * float(np.abs(np.diff(h[fx.interior], 2)).max())


def _mean_jump(dts, after_s=2000.0) -> float:
    tail = dts[np.cumsum(dts) > after_s]
    return float((np.abs(np.diff(tail)) / tail[:-1]).mean()) if tail.size > 1 else 0.0


@pytest.mark.gpu
@pytest.mark.parametrize("cadence", CADENCES)
def test_a_frequent_sync_cadence_does_not_ripple_the_reach(rbs, cadence):
    """test_clamp_ripple.py:234-258 and :293-314: interior curvature under 5 mm, the mean
    step-to-step dt change under 1 %, and no step longer than the CFL step."""
    fx, h, dts, mass, worst = _host_scheduled(rbs, cadence)
    curv, jump = _curvature_mm(fx, h), _mean_jump(dts)
    print(f"\n[clamp ripple, host loop, {cadence:g} s] {len(dts)} steps, max|d2h| {curv:.4f} mm, "
          f"dt jump mean {100 * jump:.5f} %, dt/dt_cfl max {worst:.6f}, mass {mass:.2e}")
    assert mass < MASS_GATE
    assert curv < CURVATURE_MM
    assert jump < DT_JUMP
    assert worst <= 1.0 + 1e-12


@pytest.mark.gpu
def test_the_reach_does_not_depend_on_the_cadence(rbs):
    """test_clamp_ripple.py:261-290: every cadence against an uninterrupted control."""
    fx, control, *_ = _host_scheduled(rbs, None)
    assert _curvature_mm(fx, control) < CURVATURE_MM
    for c in CADENCES:
        _, h, *_ = _host_scheduled(rbs, c)
        delta = 1000.0 * float(np.abs(h[fx.interior] - control[fx.interior]).max())
        print(f"\n[clamp ripple] {c:g} s cadence vs control: {delta:.4f} mm")
        assert delta < CURVATURE_MM


@pytest.mark.gpu
@pytest.mark.parametrize("cadence", CADENCES)
def test_device_driven_steps_fill_spans_equally(rbs, cadence):
    """The same gate on a device-driven run, whose fill_span runs on the GPU: equal to
    the host loop bit for bit, so the same curvature."""
    from rbs.driver import DeviceRun
    from rbs.scheduler import MultiRateScheduler

    fx, h_host, dts, *_ = _host_scheduled(rbs, cadence)
    fx, st, _ = _uniform_reach(rbs)
    sched = MultiRateScheduler(end_time=END_S, output_every=cadence)
    run = DeviceRun(st, "local_inertial", sched, alpha=0.7, dt_max=DT_MAX)
    steps = sum(run.run_until(seg.t_stop) for seg in sched.segments())
    h = st.h.numpy()[0].astype(np.float64)
    st.close()
    print(f"\n[clamp ripple, device-driven, {cadence:g} s] {steps} steps, max|d2h| "
          f"{_curvature_mm(fx, h):.4f} mm")
    assert steps == len(dts)
    assert np.array_equal(h, h_host)
    assert _curvature_mm(fx, h) < CURVATURE_MM
