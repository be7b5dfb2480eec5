"""Sub-grid channels on the HIP path (librbs through the C-ABI) against the oracle,
and the reference's channel gates run on the device (solver/core/test_channels.py,
validation/test_subgrid_channel.py).

Tolerance: the step's float32 arithmetic is the oracle's operation for operation, so
fields agree to the parity bound of the plain scheme (1e-5 after hundreds of steps,
in practice to the bit); zero-width arming and every tile variant are checked bit
for bit.
"""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture(scope="module")
def rbs():
    from rbs import _lib
    from rbs.core import channels, local_inertial, massbalance, sediment, state
    from rbs.processes import inflow

    _lib.load()
    assert _lib.device_count() >= 1
    return type("R", (), dict(lib=_lib, li=local_inertial, State=state.State, ch=channels,
                              sed=sediment, mb=massbalance, inflow=inflow))


def _network(ny, nx, seed):
    """A channel down the middle row and one down a column, random widths and depths;
    the floodplain a tilted random bed."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:ny, 0:nx]
    z = (0.003 * 20.0 * (nx - xx) + 0.002 * 20.0 * yy + 0.4 * rng.uniform(0, 1, (ny, nx)))
    w = np.zeros((ny, nx))
    d = np.zeros((ny, nx))
    w[ny // 2, :] = rng.uniform(4.0, 12.0, nx)
    d[ny // 2, :] = rng.uniform(0.8, 2.0, nx)
    w[:, nx // 3] = rng.uniform(3.0, 8.0, ny)
    d[:, nx // 3] = rng.uniform(0.5, 1.5, ny)
    n_ch = np.where(w > 0, rng.uniform(0.025, 0.04, (ny, nx)), 0.03)
    return z.astype(np.float32), w.astype(np.float32), d.astype(np.float32), n_ch.astype(np.float32)


def _pair(rbs, orc, z, dx, depth, manning, w, d, n_ch):
    orc.set_threads(8)
    g = rbs.State.from_bed(z, dx, depth=depth, manning=manning)
    o = orc.State(z, dx, depth=depth, manning=manning)
    rbs.ch.arm_channels(g, w, d, n_ch)
    o.arm_channels(w, d, n_ch)
    return g, o


def _dt(rbs, orc, g, o, alpha, dt_max):
    dt = rbs.li.compute_dt(g, alpha, dt_max)
    do = orc.li_compute_dt(o, alpha, dt_max)
    assert abs(dt - do) <= 1e-6 * dt, (dt, do)
    return dt


def _cmp(g, o, tol=TOL):
    chan = g.channels
    pairs = [("h", g.h.numpy(), o.h), ("qx", g.qx.numpy(), o.qx), ("qy", g.qy.numpy(), o.qy),
             ("qx_ch", chan.qx_ch.numpy(), o.qx_ch), ("qx_fp", chan.qx_fp.numpy(), o.qx_fp),
             ("qy_ch", chan.qy_ch.numpy(), o.qy_ch), ("qy_fp", chan.qy_fp.numpy(), o.qy_fp)]
    worst = {}
    for name, a, b in pairs:
        worst[name] = float(np.abs(a - b).max())
        assert worst[name] <= tol, f"max|d{name}| {worst[name]:.3e}"
    return worst


def test_storage_curve_eta_matches_the_oracle_bitwise(rbs, orc):
    """RBS_F_ETA of a channelled state is the storage curve (channels.py:59-72)."""
    z, w, d, n_ch = _network(40, 70, 1)
    h = np.random.default_rng(2).uniform(0.0, 1.2, z.shape).astype(np.float32)
    g, o = _pair(rbs, orc, z, 20.0, h, 0.05, w, d, n_ch)
    want = np.array([[orc.eta_subgrid(h[i, j], z[i, j], w[i, j], d[i, j], 20.0)
                      for j in range(z.shape[1])] for i in range(z.shape[0])], np.float32)
    assert np.array_equal(g.eta.numpy(), want)
    assert np.array_equal(g.channels.w.numpy(), w) and np.array_equal(g.channels.n.numpy(), n_ch)


def test_channel_step_parity_with_every_source_and_open_edges(rbs, orc):
    """Rain field (Kahan), infiltration, two open edges on a channel network: 400
    steps against the oracle, totals and both components."""
    ny, nx, dx = 90, 150, 20.0
    z, w, d, n_ch = _network(ny, nx, 3)
    rng = np.random.default_rng(4)
    rain = rng.uniform(1.0, 3.0, (ny, nx)).astype(np.float32) * 1e-5
    infil = rng.uniform(0.0, 4.0, (ny, nx)).astype(np.float32) * 1e-6
    g, o = _pair(rbs, orc, z, dx, 0.02, 0.06, w, d, n_ch)
    for s in (g, o):
        s.set_rain_field(rain)
        s.set_infiltration(infil)
        s.set_open_boundaries({"east": "open", "south": "open"})
        s.arm_source_compensation()
    led = rbs.mb.MassLedger.from_state(g)
    rain_sum = float(rain.astype(np.float64).sum())
    t = 0.0
    for k in range(400):
        scale = 1.0 if k < 300 else 0.0
        dt = _dt(rbs, orc, g, o, 0.7, 5.0)
        rbs.li.step(g, dt, rain_scale=scale)
        orc.li_step(o, dt, rain_scale=scale)
        if scale:
            led.add_inflow(rain_sum * dt * dx * dx)
        t += dt
    worst = _cmp(g, o)
    assert np.abs(g.channels.qx_ch.numpy()).max() > 0.0
    np.testing.assert_allclose(g.h_comp.numpy(), o.h_comp, atol=1e-7)
    assert led.record(g, t).rel_error < 1e-6
    print(f"\n[channel parity] {worst}")


@pytest.mark.parametrize("shape", [(1, 1), (1, 37), (41, 1), (2, 2), (67, 133)])
def test_odd_shapes_match(rbs, orc, shape):
    rng = np.random.default_rng(sum(shape) + 11)
    ny, nx = shape
    z = rng.uniform(0, 1, shape).astype(np.float32)
    h = rng.uniform(0.0, 1.0, shape).astype(np.float32)
    w = np.where(rng.random(shape) < 0.6, rng.uniform(0.5, 2.0, shape), 0.0).astype(np.float32)
    d = np.where(w > 0, rng.uniform(0.2, 1.0, shape), 0.0).astype(np.float32)
    g, o = _pair(rbs, orc, z, 2.0, h, 0.05, w, d, np.full(shape, 0.03, np.float32))
    for _ in range(40):
        dt = _dt(rbs, orc, g, o, 0.7, 1.0)
        rbs.li.step(g, dt)
        orc.li_step(o, dt)
    _cmp(g, o, tol=1e-6)


def test_zero_width_channels_reproduce_the_plain_kernel_bitwise(rbs):
    """test_channels.py:147-170 on the device: the channel kernel with w = 0 against the
    plain fused kernel, bit for bit, with the CFL dt of each."""
    ny, nx, dx = 64, 200, 40.0
    xx = np.broadcast_to(np.arange(nx), (ny, nx))
    bed = ((nx - 1 - xx) * dx * 0.004 + 5.0).astype(np.float32)
    bed = bed + np.random.default_rng(6).uniform(0, 0.3, (ny, nx)).astype(np.float32)
    plain = rbs.State.from_bed(bed, dx, depth=0.05, manning=0.03)
    armed = rbs.State.from_bed(bed, dx, depth=0.05, manning=0.03)
    rbs.ch.arm_channels(armed, np.zeros((ny, nx)), np.zeros((ny, nx)))
    for _ in range(60):
        dt_p = rbs.li.compute_dt(plain, 0.7, 5.0)
        dt_a = rbs.li.compute_dt(armed, 0.7, 5.0)
        assert dt_p == dt_a
        rbs.li.step(plain, dt_p, rain=1e-5)
        rbs.li.step(armed, dt_a, rain=1e-5)
    np.testing.assert_array_equal(armed.h.numpy(), plain.h.numpy())
    np.testing.assert_array_equal(armed.qx.numpy(), plain.qx.numpy())
    np.testing.assert_array_equal(armed.qy.numpy(), plain.qy.numpy())


def test_channel_variants_are_bitwise(rbs):
    """Every channel tile variant writes the same bits (fields, components, bedload)."""
    ny, nx = 173, 251
    z, w, d, n_ch = _network(ny, nx, 8)
    d50 = np.random.default_rng(9).uniform(5e-4, 4e-3, (ny, nx)).astype(np.float32)
    count = ctypes.c_int32()
    rbs.lib.call("rbs_chan_variant_count", ctypes.byref(count))
    assert count.value >= 2
    outs = []
    for v in range(count.value):
        st = rbs.State.from_bed(z, 20.0, depth=0.05, manning=0.05)
        rbs.ch.arm_channels(st, w, d, n_ch)
        rbs.sed.arm_sediment(st, d50, 0.4)
        st.set_open_boundaries({"east": "open"})
        rbs.lib.call("rbs_set_chan_variant", st._handle, v)
        for _ in range(50):
            rbs.li.step(st, rbs.li.compute_dt(st, 0.7, 3.0), rain=3e-5)
        c, s = st.channels, st.sediment
        outs.append([st.h.numpy(), st.qx.numpy(), st.qy.numpy(), c.qx_ch.numpy(),
                     c.qx_fp.numpy(), c.qy_ch.numpy(), c.qy_fp.numpy(), s.qs_int_x.numpy(),
                     s.qs_int_y.numpy()])
        st.close()
    for v, o in enumerate(outs[1:], start=1):
        for k, (a, b) in enumerate(zip(outs[0], o)):
            assert np.array_equal(a, b), f"variant {v}: field {k} differs"
    assert np.abs(outs[0][7]).max() > 0.0


def test_channel_bedload_integral_matches_the_oracle(rbs, orc):
    """accumulate_qs_*_channels (sediment.py:305-377) inside the fused step."""
    ny, nx, dx = 48, 120, 25.0
    z, w, d, n_ch = _network(ny, nx, 13)
    d50 = np.random.default_rng(14).uniform(5e-4, 3e-3, (ny, nx)).astype(np.float32)
    g, o = _pair(rbs, orc, z, dx, 0.3, 0.05, w, d, n_ch)
    rbs.sed.arm_sediment(g, d50, 0.4)
    o.arm_sediment(d50)
    for _ in range(150):
        dt = _dt(rbs, orc, g, o, 0.7, 2.0)
        rbs.li.step(g, dt, rain=5e-5)
        orc.li_step(o, dt, rain=5e-5)
    _cmp(g, o)
    sed = g.sediment
    np.testing.assert_array_equal(sed.qs_int_x.numpy(), o.qs_int_x)  # one power sequence: bitwise
    np.testing.assert_array_equal(sed.qs_int_y.numpy(), o.qs_int_y)
    assert np.abs(o.qs_int_x).max() > 0.0


def test_timestep_is_set_by_the_water_column(rbs):
    """test_channels.py:173-186 through rbs_compute_dt."""
    ny, nx, dx = 4, 8, 200.0
    w, d, h0 = 20.0, 3.0, 0.05
    bed = np.full((ny, nx), 10.0, np.float32)
    plain = rbs.State.from_bed(bed, dx, depth=h0)
    armed = rbs.State.from_bed(bed, dx, depth=h0)
    rbs.ch.arm_channels(armed, np.full((ny, nx), w), np.full((ny, nx), d))
    dt_plain = rbs.li.compute_dt(plain, 0.7, 1e6)
    dt_armed = rbs.li.compute_dt(armed, 0.7, 1e6)
    assert dt_plain == pytest.approx(0.7 * dx / math.sqrt(9.81 * h0), rel=1e-5)
    assert dt_armed == pytest.approx(0.7 * dx / math.sqrt(9.81 * h0 * dx / w), rel=1e-5)


def test_device_driven_channel_run_matches_host_loop(rbs):
    """rbs_run_until on a channelled state == the host loop, bitwise (the column CFL
    reduction and the channel kernel inside the captured graph)."""
    from rbs.driver import DeviceRun
    from rbs.scheduler import MultiRateScheduler

    z, w, d, n_ch = _network(64, 96, 21)
    sched = MultiRateScheduler(end_time=300.0, output_every=100.0, events=[200.0])
    host = rbs.State.from_bed(z, 20.0, depth=0.05, manning=0.05)
    dev = rbs.State.from_bed(z, 20.0, depth=0.05, manning=0.05)
    for s in (host, dev):
        rbs.ch.arm_channels(s, w, d, n_ch)
        s.arm_source_compensation()
    rain_m_s = 30.0 / 1000 / 3600
    n_host = 0
    for tick in sched.ticks(lambda: rbs.li.compute_dt(host, 0.7, 10.0)):
        rain = rain_m_s if tick.t0 < 200.0 - 1e-6 else 0.0
        rbs.li.step(host, tick.dt, rain=rain, t=tick.t0)
        n_host += 1
    run = DeviceRun(dev, "local_inertial", sched, alpha=0.7, dt_max=10.0, rain_m_s=rain_m_s,
                    rain_duration=200.0)
    n_dev = sum(run.run_until(s.t_stop) for s in sched.segments())
    assert n_dev == n_host
    assert np.array_equal(host.h.numpy(), dev.h.numpy())
    assert np.array_equal(host.channels.qx_ch.numpy(), dev.channels.qx_ch.numpy())


def test_subgrid_channel_reaches_manning_normal_depth_on_the_device(rbs):
    """validation/test_subgrid_channel.py:101-140 through the HIP path."""
    from test_oracle_channels import normal_depth

    nx, dx = 100, 50.0
    slope, n_ch, width, bankfull, q = 0.001, 0.03, 20.0, 4.0, 30.0
    xx = np.arange(nx)[None, :]
    bed = ((nx - 1 - xx) * dx * slope + 5.0).astype(np.float32)
    st = rbs.State.from_bed(bed, dx, depth=0.0, manning=0.05)
    rbs.ch.arm_channels(st, np.full((1, nx), width), np.full((1, nx), bankfull),
                        np.full((1, nx), n_ch))
    st.set_open_boundaries({"east": "open"})
    inj = rbs.inflow.InflowInjector([rbs.inflow.Inflow((0, 0), [(0.0, q), (1.0e9, q)])], st)
    led = rbs.mb.MassLedger.from_state(st)
    t = 0.0
    for _ in range(3000):
        dt = rbs.li.compute_dt(st, 0.7, 10.0)
        led.add_inflow(inj.apply(st, t, dt))
        rbs.li.step(st, dt, t=t)
        t += dt
    rec = led.record(st, t)
    column = st.h.numpy()[0] * dx / width
    median = float(np.median(column[10:nx - 10]))
    h_n = normal_depth(q, width, slope, n_ch)
    print(f"\n[subgrid] normal={h_n:.3f} m modelled={median:.3f} m mass={rec.rel_error:.2e}")
    assert rec.rel_error < 1e-6
    assert median < bankfull
    assert median == pytest.approx(h_n, rel=0.05)


def test_a_batch_refuses_a_channelled_state(rbs):
    from rbs.driver import DeviceRun
    from rbs.ensemble import Batch
    from rbs.scheduler import MultiRateScheduler

    z, w, d, n_ch = _network(16, 16, 5)
    st = rbs.State.from_bed(z, 20.0, depth=0.1)
    rbs.ch.arm_channels(st, w, d, n_ch)
    run = DeviceRun(st, "local_inertial", MultiRateScheduler(end_time=100.0, output_every=50.0),
                    alpha=0.7, dt_max=5.0)
    b = Batch(capacity=2)
    try:
        with pytest.raises(Exception, match="channels"):
            b.add(run, 100.0)
    finally:
        b.close()
