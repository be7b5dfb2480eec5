"""HLLC HIP path (two fused SSP-RK2 stage kernels) against the oracle.

Same tolerance as the local-inertial path: max|dh| and max|d(hu, hv)| <= 1e-5
after the stated number of steps with identical dt sequences, against the oracle's
default mode (the correctly rounded float32-exponent powers the kernels evaluate)
and its libm mode (glibc powf, as the reference's Warp-CPU runs evaluated).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def rbs():
    from rbs import _lib
    from rbs.core import hllc, massbalance, state

    _lib.load()
    return type("R", (), dict(lib=_lib, hl=hllc, State=state.State,
                              MassLedger=massbalance.MassLedger))


def _bumpy(ny, nx, dx):
    yy, xx = np.mgrid[0:ny, 0:nx]
    return (0.5 * np.sin(xx * 0.2) + 0.002 * xx * dx + 0.3 * np.cos(yy * 0.3)).astype(np.float32)


def _cmp(g, o, tol=TOL):
    d = [np.abs(g.h.numpy() - o.h).max(), np.abs(g.hu.numpy() - o.hu).max(),
         np.abs(g.hv.numpy() - o.hv).max()]
    assert max(d) <= tol, f"max|dU| {d}"
    return d


def _run_pair(rbs, orc, g, o, steps, alpha=0.45, dt_max=5.0, rain=0.0, t0=0.0):
    t = t0
    for _ in range(steps):
        dt = rbs.hl.compute_dt(g, alpha, dt_max)
        do = orc.hllc_compute_dt(o, alpha, dt_max)
        assert abs(dt - do) <= 1e-6 * dt, (dt, do)
        rbs.hl.step(g, dt, rain=rain, t=t)
        orc.hllc_step(o, dt, rain=rain, t=t)
        t += dt
    return t


def test_moving_water_on_a_bumpy_bed_closed_walls(rbs, orc):
    orc.set_threads(8)
    z = _bumpy(70, 150, 10.0)
    yy, xx = np.mgrid[0:70, 0:150]
    d0 = np.maximum(0.5 + 0.4 * np.sin(xx * 0.13 + yy * 0.21) - 0.2 * z, 0.0).astype(np.float32)
    g = rbs.State.from_bed(z, 10.0, depth=d0, manning=0.03)
    o = orc.State(z, 10.0, depth=d0, manning=0.03)
    _run_pair(rbs, orc, g, o, 200)
    print("\n[hllc closed] max|dU|", _cmp(g, o))


def test_open_and_stage_edges_rain_infiltration(rbs, orc):
    """Open west/north, time-varying stage on the east, wall south; Kahan rain field,
    infiltration; ledger totals agree with the oracle's loss_cum."""
    rng = np.random.default_rng(4)
    ny, nx = 48, 96
    yy, xx = np.mgrid[0:ny, 0:nx]
    z = (0.01 * (nx - xx) + 0.2 * rng.uniform(0, 1, (ny, nx))).astype(np.float32)
    d0 = np.maximum(0.6 - z, 0.0).astype(np.float32)
    rain = rng.uniform(1, 3, (ny, nx)).astype(np.float32) * 1e-5
    infil = rng.uniform(0, 5, (ny, nx)).astype(np.float32) * 1e-6
    curve = [(0.0, 0.5), (200.0, 0.9), (400.0, 0.4)]
    g = rbs.State.from_bed(z, 5.0, depth=d0, manning=0.035)
    o = orc.State(z, 5.0, depth=d0, manning=0.035)
    o_libm = orc.State(z, 5.0, depth=d0, manning=0.035)
    for s in (g, o, o_libm):
        s.set_open_boundaries({"west": "open", "north": "open", "east": "fixed_stage"},
                              {"east": curve})
        s.set_rain_field(rain)
        s.set_infiltration(infil)
        s.arm_source_compensation()
    led_g, led_o = rbs.MassLedger.from_state(g), orc.MassLedger.from_state(o)
    t = 0.0
    for k in range(300):
        dt = rbs.hl.compute_dt(g, 0.45, 2.0)
        assert abs(dt - orc.hllc_compute_dt(o, 0.45, 2.0)) <= 1e-6 * dt
        scale = 1.0 if k < 200 else 0.0
        rbs.hl.step(g, dt, rain_scale=scale, t=t)
        orc.hllc_step(o, dt, rain_scale=scale, t=t)
        with orc.pow_mode("libm"):
            orc.hllc_step(o_libm, dt, rain_scale=scale, t=t)
        if scale:
            s = float(rain.astype(np.float64).sum()) * dt * 25.0
            led_g.add_inflow(s)
            led_o.add_inflow(s)
        t += dt
    d = _cmp(g, o)
    d_libm = _cmp(g, o_libm)
    rg, ro = led_g.record(g, t), led_o.record(o, t)
    assert rg.outflow_cum == pytest.approx(ro["outflow_cum"], rel=1e-6, abs=1e-6)
    assert rg.rel_error < 1e-6
    print(f"\n[hllc edges] max|dU| {d} (libm oracle {d_libm}) outflow {rg.outflow_cum:.4f} "
          f"mass {rg.rel_error:.2e}")


@pytest.mark.parametrize("compensated", [False, True])
def test_uniform_rain(rbs, orc, compensated):
    """hllc.py:1351-1356: uniform rain, added plainly (_add_uniform_rain, a24) or by
    the Kahan kernel when the source compensation is armed; h, hu, hv and the Kahan
    debt against the oracle, with the storm stopping half way."""
    rng = np.random.default_rng(14)
    ny, nx = 40, 72
    yy, xx = np.mgrid[0:ny, 0:nx]
    z = (0.008 * (nx - xx) + 0.15 * rng.uniform(0, 1, (ny, nx))).astype(np.float32)
    d0 = np.maximum(0.4 - z, 0.0).astype(np.float32)
    g = rbs.State.from_bed(z, 5.0, depth=d0, manning=0.03)
    o = orc.State(z, 5.0, depth=d0, manning=0.03)
    for s in (g, o):
        s.set_open_boundaries({"west": "open"})
        if compensated:
            s.arm_source_compensation()
    rain = 30.0 / 1000 / 3600
    t = _run_pair(rbs, orc, g, o, 150, dt_max=2.0, rain=rain)
    _run_pair(rbs, orc, g, o, 150, dt_max=2.0, rain=0.0, t0=t)
    d = _cmp(g, o)
    if compensated:
        np.testing.assert_array_equal(g.h_comp.numpy(), o.h_comp)
        assert np.any(o.h_comp != 0.0)
    assert float(g.h.numpy().astype(np.float64).sum()) > float(d0.astype(np.float64).sum())
    print(f"\n[hllc uniform rain, compensated={compensated}] max|dU| {d}")


def test_wet_bed_stoker_on_gpu(rbs, orc):
    """validation/test_dam_break.py:93-133 for HLLC: nRMSE < 0.03, front < 0.05."""
    nx, dx, t_end = 800, 0.5, 8.0
    x = (np.arange(nx) + 0.5 - nx / 2.0) * dx
    g = rbs.State.from_bed(np.zeros((1, nx)), dx, depth=np.where(x < 0, 1.0, 0.3).reshape(1, nx),
                           manning=0.01)
    led = rbs.MassLedger.from_state(g)
    t = 0.0
    while t < t_end - 1e-9:
        dt = min(rbs.hl.compute_dt(g, 0.45, t_end), t_end - t)
        rbs.hl.step(g, dt)
        t += dt
    rec = led.record(g, t)
    h = g.h.numpy()[0].astype(np.float64)
    h_ref, _ = orc.stoker(x, t_end, 1.0, 0.3)
    inter = np.abs(x) < 0.9 * 200
    nrmse = math.sqrt(np.mean((h[inter] - h_ref[inter]) ** 2)) / 0.7
    lvl = 0.5 * (0.3 + orc.stoker_star_depth(1.0, 0.3))
    fr = orc.stoker_front_position(t_end, 1.0, 0.3)
    front = abs(x[h > lvl].max() - fr) / fr
    assert nrmse < 0.03 and front < 0.05 and rec.rel_error < 1e-6
    assert round(nrmse, 4) == 0.0076  # M4-fidelity-step.md:197


def test_lake_at_rest_on_gpu(rbs):
    """validation/test_hllc_2d.py:36-55 (gate 1e-4); the achieved residual is printed."""
    z = _bumpy(30, 40, 10.0)
    eta0 = float(z.max()) + 2.0
    g = rbs.State.from_bed(z, 10.0, depth=np.maximum(eta0 - z, 0.0), manning=0.03)
    for _ in range(200):
        rbs.hl.step(g, rbs.hl.compute_dt(g, 0.45, 5.0))
    u, v = g.velocities_numpy()
    worst = max(np.abs(u).max(), np.abs(v).max())
    print(f"\n[hllc lake at rest] max|u,v| = {worst:.2e} m/s")
    assert worst < 1e-4


def test_device_driven_hllc_matches_host_loop(rbs):
    from rbs.driver import DeviceRun
    from rbs.scheduler import MultiRateScheduler

    z = _bumpy(40, 64, 5.0)
    yy, xx = np.mgrid[0:40, 0:64]
    d0 = np.maximum(0.8 + 0.3 * np.sin(xx * 0.3) - 0.3 * z, 0).astype(np.float32)
    curve = [(0.0, 0.9), (60.0, 1.2)]
    sched = MultiRateScheduler(end_time=120.0, output_every=40.0, events=[60.0])
    host = rbs.State.from_bed(z, 5.0, depth=d0, manning=0.03)
    dev = rbs.State.from_bed(z, 5.0, depth=d0, manning=0.03)
    for s in (host, dev):
        s.set_open_boundaries({"east": "fixed_stage", "west": "open"}, {"east": curve})
    n_host = 0
    for tick in sched.ticks(lambda: rbs.hl.compute_dt(host, 0.45, 10.0)):
        rbs.hl.step(host, tick.dt, t=tick.t0)
        n_host += 1
    run = DeviceRun(dev, "hllc_fv", sched, alpha=0.45, dt_max=10.0)
    n_dev = sum(run.run_until(seg.t_stop) for seg in sched.segments())
    assert n_dev == n_host
    assert np.array_equal(host.h.numpy(), dev.h.numpy())
    assert np.array_equal(host.hu.numpy(), dev.hu.numpy())
    assert host.loss_volume() == pytest.approx(dev.loss_volume(), rel=1e-12)


def test_tile_variants_are_bitwise(rbs):
    """Every HLLC tile variant writes the same bits (open, stage and wall edges, rain
    field with Kahan debt, infiltration)."""
    import ctypes

    rng = np.random.default_rng(8)
    ny, nx = 83, 141  # ragged against every tile shape
    yy, xx = np.mgrid[0:ny, 0:nx]
    z = (0.01 * (nx - xx) + 0.2 * rng.uniform(0, 1, (ny, nx))).astype(np.float32)
    d0 = np.maximum(0.6 - z + 0.1 * np.sin(0.3 * yy), 0.0).astype(np.float32)
    rain = rng.uniform(1, 3, (ny, nx)).astype(np.float32) * 1e-5
    infil = rng.uniform(0, 5, (ny, nx)).astype(np.float32) * 1e-6
    count = ctypes.c_int32()
    rbs.lib.call("rbs_hllc_variant_count", ctypes.byref(count))
    outs = []
    for v in range(count.value):
        g = rbs.State.from_bed(z, 5.0, depth=d0, manning=0.035)
        g.set_open_boundaries({"west": "open", "east": "fixed_stage"}, {"east": [(0.0, 0.55)]})
        g.set_rain_field(rain)
        g.set_infiltration(infil)
        g.arm_source_compensation()
        rbs.lib.call("rbs_set_hllc_variant", g._handle, v)
        t = 0.0
        for _ in range(40):
            dt = rbs.hl.compute_dt(g, 0.45, 2.0)
            rbs.hl.step(g, dt, rain_scale=1.0, t=t)
            t += dt
        outs.append((g.h.numpy(), g.hu.numpy(), g.hv.numpy(), g.h_comp.numpy()))
        g.close()
    for v, o in enumerate(outs[1:], start=1):
        for name, a, b in zip(("h", "hu", "hv", "h_comp"), outs[0], o):
            assert np.array_equal(a, b), f"variant {v}: {name} differs"


def test_variants_are_bitwise_with_bedload_and_uniform_rain(rbs):
    """Every variant, tiles and strips, with the bedload integral armed, uniform rain on
    a Kahan debt, and the other edge types (closed west, open east, fixed-stage north,
    open south): the same h, hu, hv, debts and transport integrals."""
    import ctypes

    from rbs.core import sediment

    rng = np.random.default_rng(11)
    ny, nx = 97, 131
    yy, xx = np.mgrid[0:ny, 0:nx]
    z = (0.008 * (nx - xx) + 0.004 * yy + 0.15 * rng.uniform(0, 1, (ny, nx))).astype(np.float32)
    d0 = np.maximum(0.7 - z + 0.15 * np.cos(0.2 * xx), 0.0).astype(np.float32)
    d50 = rng.uniform(5e-4, 3e-3, (ny, nx)).astype(np.float32)
    count = ctypes.c_int32()
    rbs.lib.call("rbs_hllc_variant_count", ctypes.byref(count))
    outs, losses = [], []
    for v in range(count.value):
        g = rbs.State.from_bed(z, 4.0, depth=d0, manning=0.03)
        g.set_open_boundaries({"west": "closed", "east": "open", "north": "fixed_stage",
                               "south": "open"}, {"north": [(0.0, 0.9), (600.0, 1.1)]})
        g.arm_source_compensation()
        sediment.arm_sediment(g, d50, 0.4)
        rbs.lib.call("rbs_set_hllc_variant", g._handle, v)
        t = 0.0
        for _ in range(60):
            dt = rbs.hl.compute_dt(g, 0.45, 2.0)
            rbs.hl.step(g, dt, rain=2e-5, t=t)
            t += dt
        sed = g.sediment
        outs.append((g.h.numpy(), g.hu.numpy(), g.hv.numpy(), g.h_comp.numpy(),
                     sed.qs_int_x.numpy(), sed.qs_int_y.numpy()))
        losses.append(g.loss_volume())
        g.close()
    assert np.abs(outs[0][4]).max() > 0.0 and np.abs(outs[0][1]).max() > 0.0
    names = ("h", "hu", "hv", "h_comp", "qs_int_x", "qs_int_y")
    for v, o in enumerate(outs[1:], start=1):
        for name, a, b in zip(names, outs[0], o):
            assert np.array_equal(a, b), f"variant {v}: {name} differs"
        # the edge bank is a float64 sum whose order follows the tiling
        assert losses[v] == pytest.approx(losses[0], rel=1e-12, abs=1e-9)


def test_dry_floodplain_variants_and_oracle_are_bitwise(rbs, orc):
    """A channel in a dry floodplain without rain, so that whole rows of a strip stay
    exactly dry (h = +0) and its iterations take the dry path; a patch of the bed at
    -0.0; a fixed-stage edge (north) that can wet a dry cell, and a closed and two open
    edges.  Every variant writes the same bits, and the default equals the oracle."""
    import ctypes

    rng = np.random.default_rng(21)
    ny, nx = 150, 301
    yy, xx = np.mgrid[0:ny, 0:nx]
    z = (0.004 * np.abs(xx - 150) + 0.0005 * (ny - yy) + 0.02 * rng.uniform(0, 1, (ny, nx)))
    z = z.astype(np.float32)
    z[100:112, 240:262] = np.float32(-0.0)
    assert np.signbit(z[105, 250])
    d0 = np.where(np.abs(xx - 150) < 12, 0.6, 0.0).astype(np.float32)
    bnd = {"west": "closed", "east": "open", "north": "fixed_stage", "south": "open"}
    curve = {"north": [(0.0, 0.45), (60.0, 0.8)]}

    def run(v=None):
        g = rbs.State.from_bed(z, 5.0, depth=d0, manning=0.03)
        g.set_open_boundaries(bnd, curve)
        if v is not None:
            rbs.lib.call("rbs_set_hllc_variant", g._handle, v)
        t = 0.0
        for _ in range(80):
            dt = rbs.hl.compute_dt(g, 0.45, 2.0)
            rbs.hl.step(g, dt, t=t)
            t += dt
        out = (g.h.numpy(), g.hu.numpy(), g.hv.numpy())
        g.close()
        return out, t

    count = ctypes.c_int32()
    rbs.lib.call("rbs_hllc_variant_count", ctypes.byref(count))
    outs = [run(v)[0] for v in range(count.value)]
    for v, o in enumerate(outs[1:], start=1):
        for name, a, b in zip(("h", "hu", "hv"), outs[0], o):
            assert np.array_equal(a, b), f"variant {v}: {name} differs"
    # the run keeps a dry floodplain and wets the north edge through the stage
    h = outs[0][0]
    assert (h == 0.0).mean() > 0.5 and h[0].max() > 0.0
    # the default variant against the oracle, step for step
    o = orc.State(z, 5.0, depth=d0, manning=0.03)
    o.arm_hllc()
    o.set_open_boundaries(bnd, curve)
    g = rbs.State.from_bed(z, 5.0, depth=d0, manning=0.03)
    g.set_open_boundaries(bnd, curve)
    t = 0.0
    for _ in range(80):
        dt = rbs.hl.compute_dt(g, 0.45, 2.0)
        assert abs(dt - orc.hllc_compute_dt(o, 0.45, 2.0)) <= 1e-6 * dt
        rbs.hl.step(g, dt, t=t)
        orc.hllc_step(o, dt, t=t)
        t += dt
    for name, a, b in (("h", g.h.numpy(), o.h), ("hu", g.hu.numpy(), o.hu), ("hv", g.hv.numpy(), o.hv)):
        assert np.array_equal(a, b), f"oracle: {name} differs"
    g.close()


def test_dry_blocks_wake_for_sources_and_uploads(rbs):
    """Blocks that were dry and skipped their stages come back when water reaches them:
    a point source that starts into a dry field, an upload of depth into a dry area
    and an upload of momentum onto dry cells (which the next step zeroes) mid-run.
    The default strips, which skip dry blocks, against the LDS tiles, which compute
    every cell: the same bits, step for step at the end of each phase."""
    from rbs.processes.inflow import Inflow, InflowInjector

    rng = np.random.default_rng(5)
    ny, nx = 130, 257
    yy, xx = np.mgrid[0:ny, 0:nx]
    z = (0.003 * np.abs(xx - 60) + 0.02 * rng.uniform(0, 1, (ny, nx))).astype(np.float32)
    d0 = np.where(np.abs(xx - 60) < 6, 0.5, 0.0).astype(np.float32)
    blob = np.zeros((ny, nx), np.float32)
    blob[90:100, 200:215] = 0.4
    mom = np.zeros((ny, nx), np.float32)
    mom[20:40, 150:250] = 0.3

    def run(variant):
        g = rbs.State.from_bed(z, 5.0, depth=d0, manning=0.03)
        g.set_open_boundaries({"east": "open"})
        if variant is not None:
            rbs.lib.call("rbs_set_hllc_variant", g._handle, variant)
        inj = InflowInjector([Inflow((60, 180), [(0.0, 0.0), (30.0, 0.0), (31.0, 4.0),
                                                 (1e9, 4.0)])], g)
        snaps, t = [], 0.0
        for k in range(90):
            if k == 40:
                g.h.assign(g.h.numpy() + blob)
                g.hu.assign(np.where(g.h.numpy() == 0.0, mom, g.hu.numpy()))
            dt = rbs.hl.compute_dt(g, 0.45, 2.0)
            inj.apply(g, t, dt)
            rbs.hl.step(g, dt, t=t)
            t += dt
            if k in (39, 41, 89):
                snaps.append((g.h.numpy(), g.hu.numpy(), g.hv.numpy()))
        g.close()
        return snaps, t

    strips, t_s = run(None)
    tiles, t_t = run(4)
    assert t_s == t_t
    for phase, (a, b) in enumerate(zip(strips, tiles)):
        for name, x, y in zip(("h", "hu", "hv"), a, b):
            assert np.array_equal(x, y), f"phase {phase}: {name} differs"
    h_end = strips[-1][0]
    assert h_end[60, 180] > 0.0 and h_end[95, 207] > 0.0  # both new waters are there
    assert (h_end == 0.0).mean() > 0.3  # and much of the field is still dry


def test_many_blocks_fold_in_two_levels(rbs):
    """A 2048^2 HLLC state has 4608 strip blocks, past the point where the block
    partials are folded in two levels: the ledger's volume equals the float64 sum of
    the depths, and the wave-speed maximum the float32 formula over the whole grid."""
    rng = np.random.default_rng(3)
    n, dx = 2048, 5.0
    z = (0.5 * rng.uniform(0, 1, (n, n))).astype(np.float32)
    h0 = np.maximum(1.2 - z + 0.2 * rng.uniform(0, 1, (n, n)), 0.0).astype(np.float32)
    g = rbs.State.from_bed(z, dx, depth=h0, manning=0.03)
    for _ in range(3):
        rbs.hl.step(g, rbs.hl.compute_dt(g, 0.45, 2.0))
    led = g.ledger()
    h, hu, hv = g.h.numpy(), g.hu.numpy(), g.hv.numpy()
    vol = float(h.astype(np.float64).sum()) * dx * dx
    assert led.volume == pytest.approx(vol, rel=1e-12)
    c = np.sqrt(np.float32(9.81) * np.maximum(h, np.float32(0.0)))
    wet = h > np.float32(1e-3)
    hs = np.where(wet, h, np.float32(1.0))
    sp = np.where(wet, np.maximum(np.abs(hu) / hs, np.abs(hv) / hs) + c, c)
    assert np.float32(led.speed_max) == sp.max()
    g.close()


@pytest.mark.parametrize("name", ["flow_over_bump", "wet_dam_break_bumpy_slope",
                                  "lake_over_two_bumps"])
def test_wet_flows_match_the_reference_1d_solver(rbs, orc, name):
    """The golden wet cases of the reference's float64 1-D solver (validation/hllc1d.py,
    tests/golden/hllc1d_vectors.npz) on the HIP path: within the oracle's tolerance of
    them (5e-6 m, 3e-5 m/s), and bitwise the oracle."""
    from test_oracle_hllc import _vectors, run_1d_case

    def make_g(z, dx, h0, hu0, n):
        st = rbs.State.from_bed(z, dx, depth=h0, manning=n)
        st.set_open_boundaries({"west": "open", "east": "open"})
        rbs.hl.arm_hllc(st)
        st.hu.assign(hu0)
        return st

    def make_o(z, dx, h0, hu0, n):
        st = orc.State(z, dx, depth=h0, manning=n)
        st.arm_hllc()
        st.hu[...] = hu0
        st.set_open_boundaries({"west": "open", "east": "open"})
        return st

    d = _vectors()
    g, h_ref, u_ref = run_1d_case(lambda s, dt, t: rbs.hl.step(s, dt, t=t),
                                  lambda s: rbs.hl.compute_dt(s, 0.45, 1e9), make_g, d, name)
    o, _, _ = run_1d_case(lambda s, dt, t: orc.hllc_step(s, dt, t=t),
                          lambda s: orc.hllc_compute_dt(s, 0.45, 1e9), make_o, d, name)
    h = g.h.numpy()[0].astype(np.float64)
    u = np.where(h > 1e-3, g.hu.numpy()[0] / np.maximum(h, 1e-3), 0.0)
    print(f"\n[hllc vs 1-D reference, {name}] max|dh| {np.abs(h - h_ref).max():.2e} m, "
          f"max|du| {np.abs(u - u_ref).max():.2e} m/s")
    assert np.array_equal(g.h.numpy(), o.h) and np.array_equal(g.hu.numpy(), o.hu)
    assert np.abs(h - h_ref).max() < 5e-6
    assert np.abs(u - u_ref).max() < 3e-5
