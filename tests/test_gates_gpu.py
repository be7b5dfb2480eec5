"""The reference's validation gates on the HIP path, and the HLLC valley workload's
own checks:

* open edge drains and stays non-negative, Manning normal depth (local inertial):
  validation/test_channel_flow.py:37-108;
* Manning normal depth within 2 % (HLLC): validation/test_hllc_boundaries.py:221-265;
* the dry-bed Ritter dam break, both schemes: validation/test_dam_break.py:142-175;
* the shoreline lake at rest on a bumpy bed with dry islands:
  validation/test_hllc_wetdry.py:49-108;
* the datum shift restoring a lake at rest at altitude, and depth independent of the
  datum: solver/core/test_datum.py:60-120;
* the valley: a datum-shifted lake at rest on a 1024^2 crop of the 8192^2 bed, and a
  256^2 crop with the inflow hydrograph and the fixed-stage outlet against the oracle.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MASS_GATE = 1e-6
H_DRY = 1e-3


@pytest.fixture(scope="module")
def rbs():
    from rbs import _lib
    from rbs.core import datum, hllc, local_inertial, massbalance, state
    from rbs.processes import inflow

    _lib.load()
    return type("R", (), dict(lib=_lib, li=local_inertial, hl=hllc, State=state.State,
                              mb=massbalance, inflow=inflow, datum=datum))


def _east_open(st):
    st.set_open_boundaries({"east": "open"})


def test_li_open_edge_drains_and_stays_non_negative(rbs):
    """test_channel_flow.py:37-65: a 2 % basin draining through its east edge."""
    ny, nx, dx = 20, 40, 10.0
    bed = np.broadcast_to((nx - 1 - np.arange(nx)) * dx * 0.02, (ny, nx)).astype(np.float32)
    st = rbs.State.from_bed(bed, dx, depth=0.5, manning=0.03)
    _east_open(st)
    led = rbs.mb.MassLedger.from_state(st)
    t = 0.0
    for _ in range(2000):
        dt = rbs.li.compute_dt(st, 0.7, 5.0)
        rbs.li.step(st, dt)
        t += dt
    rec = led.record(st, t)
    h = st.h.numpy()
    assert np.isfinite(h).all() and h.min() >= -1e-6
    assert led.max_rel_error < MASS_GATE
    assert rec.outflow_cum > 0.5 * led.v0 and rec.volume < 0.5 * led.v0


def _channel(rbs, scheme: str, nx: int, slope: float, steps: int):
    dx, n, discharge = 10.0, 0.03, 2.0
    bed = ((nx - 1 - np.arange(nx)) * dx * slope).reshape(1, nx).astype(np.float32)
    st = rbs.State.from_bed(bed, dx, depth=0.0, manning=n)
    _east_open(st)
    inj = rbs.inflow.InflowInjector(
        [rbs.inflow.Inflow((0, 0), [(0.0, discharge), (1.0e9, discharge)])], st)
    led = rbs.mb.MassLedger.from_state(st)
    mod, alpha = (rbs.li, 0.7) if scheme == "li" else (rbs.hl, 0.45)
    t = 0.0
    for _ in range(steps):
        dt = mod.compute_dt(st, alpha, 10.0)
        led.add_inflow(inj.apply(st, t, dt))
        mod.step(st, dt, t=t)
        t += dt
    rec = led.record(st, t)
    h_n = (discharge / dx * n / math.sqrt(slope)) ** 0.6
    return st.h.numpy()[0], h_n, led, rec


def test_li_channel_reaches_manning_normal_depth(rbs):
    """test_channel_flow.py:68-108: the interior median within [0.5, 2] h_n."""
    h, h_n, led, rec = _channel(rbs, "li", 80, 0.001, 4000)
    median = float(np.median(h[5:75]))
    print(f"\n[LI channel] h_n {h_n:.3f} m, interior median {median:.3f} m, "
          f"mass {rec.rel_error:.2e}")
    assert np.isfinite(h).all() and led.max_rel_error < MASS_GATE and rec.outflow_cum > 0.0
    assert 0.5 * h_n < median < 2.0 * h_n


def test_hllc_channel_reaches_manning_normal_depth(rbs):
    """test_hllc_boundaries.py:221-265: transcritical (Fr 1.1), within 2 % of h_n."""
    h, h_n, led, rec = _channel(rbs, "hllc", 120, 0.02, 8000)
    median = float(np.median(h[10:110]))
    print(f"\n[HLLC channel] h_n {h_n:.4f} m, interior median {median:.4f} m "
          f"({median / h_n:.4f}x), mass {rec.rel_error:.2e}")
    assert np.isfinite(h).all() and h.min() > H_DRY
    assert led.max_rel_error < MASS_GATE and rec.outflow_cum > 0.0
    assert abs(median - h_n) < 0.02 * h_n


@pytest.mark.parametrize("scheme,alpha", [("li", 0.7), ("hllc", 0.45)])
def test_dry_bed_ritter(rbs, orc, scheme, alpha):
    """test_dam_break.py:142-175: h 1 m onto 1 mm, 800 cells of 0.5 m, n 0.02, 6 s: the
    mass gate, and the Ritter shape as a diagnostic with a blow-up ceiling."""
    nx, dx, t_end = 800, 0.5, 6.0
    x = (np.arange(nx) + 0.5 - nx / 2.0) * dx
    st = rbs.State.from_bed(np.zeros((1, nx)), dx,
                            depth=np.where(x < 0, 1.0, 1e-3).reshape(1, nx), manning=0.02)
    led = rbs.mb.MassLedger.from_state(st)
    mod = rbs.li if scheme == "li" else rbs.hl
    t = 0.0
    while t < t_end - 1e-9:
        dt = min(mod.compute_dt(st, alpha, t_end), t_end - t)
        mod.step(st, dt)
        t += dt
    led.record(st, t)
    h = st.h.numpy()[0].astype(np.float64)
    h_ref, _ = orc.ritter(x, t_end, 1.0)
    inter = np.abs(x) < 0.9 * 200.0
    nrmse = math.sqrt(np.mean((h[inter] - h_ref[inter]) ** 2))
    print(f"\n[{scheme} dry-bed Ritter] nRMSE {nrmse:.4f}, mass {led.max_rel_error:.2e}")
    assert led.max_rel_error < MASS_GATE and np.isfinite(h).all()
    assert nrmse < 0.5


def test_hllc_shoreline_lake_at_rest(rbs):
    """test_hllc_wetdry.py:49-108: a pool with three bumps piercing it as dry islands;
    velocities stay at the float32 floor, the surface does not drift, no water climbs
    onto the islands, no mass is lost."""
    ny = nx = 50
    yy, xx = np.mgrid[0:ny, 0:nx].astype(float)
    z = 0.012 * ((xx - 24.5) ** 2 + (yy - 24.5) ** 2)
    for cx, cy, a, w in ((16, 27, 2.6, 22.0), (33, 21, 3.2, 26.0), (27, 34, 2.2, 18.0)):
        z = z + a * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / w)
    z = z.astype(np.float32)
    h0 = np.maximum(3.0 - z, 0.0).astype(np.float32)
    wet, dry = h0 >= H_DRY, h0 < H_DRY
    assert not (wet[0].any() or wet[-1].any() or wet[:, 0].any() or wet[:, -1].any())
    shore = dry[1:-1, 1:-1] & (wet[:-2, 1:-1] | wet[2:, 1:-1] | wet[1:-1, :-2] | wet[1:-1, 2:])
    assert shore.sum() > 50
    st = rbs.State.from_bed(z, 5.0, depth=h0, manning=0.03)
    v0 = float(st.h.numpy().sum())
    peak = 0.0
    for _ in range(100):
        rbs.hl.step(st, rbs.hl.compute_dt(st, 0.45, 5.0))
        u, v = st.velocities_numpy()
        peak = max(peak, float(np.abs(u).max()), float(np.abs(v).max()))
    h = st.h.numpy()
    drift = float(np.abs((h + z)[wet] - 3.0).max())
    leak = float(h[dry].max())
    print(f"\n[shoreline lake] peak |u,v| {peak:.3e}, surface drift {drift:.3e}, "
          f"dry leak {leak:.3e}, dV {abs(float(h.sum()) - v0):.3e}")
    assert np.isfinite(h).all() and h.min() >= 0.0
    assert peak < 5e-5 and drift < 1e-4 and leak < H_DRY
    assert abs(float(h.sum()) - v0) < 1e-6 * v0


def _bumpy(ny, nx):
    yy, xx = np.mgrid[0:ny, 0:nx].astype(np.float64)
    return (0.05 * xx + 0.03 * yy + 0.8 * np.sin(0.7 * xx) * np.cos(0.5 * yy)
            + 0.4 * np.sin(0.3 * xx)).astype(np.float32)


def _lake_peak_speed(rbs, bed, surface, steps=60, manning=0.0):
    h0 = np.maximum(surface - bed.astype(np.float64), 0.0).astype(np.float32)
    st = rbs.State.from_bed(bed, 10.0, depth=h0, manning=manning)
    worst = 0.0
    for _ in range(steps):
        rbs.hl.step(st, rbs.hl.compute_dt(st, 0.45, 5.0))
        u, v = st.velocities_numpy()
        worst = max(worst, float(np.abs(u).max()), float(np.abs(v).max()))
    st.close()
    return worst


@pytest.mark.parametrize("datum", [5000.0, 500.0])
def test_the_datum_shift_restores_a_lake_at_rest_at_altitude(rbs, datum):
    """test_datum.py:72-96: the same lake with the bed 500 or 5000 m up, shifted back to
    0 and not: shifted stays at rest (1e-4), unshifted is at least 10x worse."""
    bed0 = _bumpy(40, 40)
    surface0 = float(bed0.max()) + 0.5
    raw_bed = bed0 + np.float32(datum)
    z_ref = rbs.datum.resolve_datum(datum, raw_bed)
    shifted = _lake_peak_speed(rbs, rbs.datum.shift_bed(raw_bed, z_ref), surface0)
    raw = _lake_peak_speed(rbs, raw_bed, surface0 + datum)
    print(f"\n[datum {datum:g} m] max|u| raw {raw:.3e} m/s, shifted {shifted:.3e} m/s")
    assert shifted < 1e-4
    assert raw > 10.0 * shifted


def test_depth_does_not_depend_on_the_datum(rbs):
    """test_datum.py:99-120: a flow on the bed and on the bed lifted 300 m and shifted
    back agree to a fraction of a millimetre (the bed requantised by ULP(300 m))."""
    bed0 = _bumpy(24, 24)
    surface0 = float(bed0.max()) + 0.3

    def final(bed):
        h0 = np.maximum(surface0 - bed0.astype(np.float64), 0.0).astype(np.float32)
        st = rbs.State.from_bed(bed, 10.0, depth=h0, manning=0.03)
        for _ in range(20):
            rbs.hl.step(st, rbs.hl.compute_dt(st, 0.45, 2.0))
        out = st.h.numpy()
        st.close()
        return out

    a = final(bed0)
    b = final(rbs.datum.shift_bed(bed0 + np.float32(300.0), 300.0))
    assert float(np.abs(a - b).max()) < 1e-4


@pytest.fixture(scope="module")
def valley():
    from rbs import terrain

    return terrain.valley(8192, dx=5.0, seed=20261204)


def test_valley_lake_at_rest_with_the_datum_shift(rbs, valley):
    """Flat water 10 m above the lowest bed of the central 1024^2 crop of the valley,
    closed walls, 1000 steps, with the datum shift ('auto': floor of the lowest bed) and
    without: the reference's 1e-4 gate on the shifted lake; the 1e-6 target of the
    workload is reported beside it."""
    n, side = 8192, 1024
    r0 = (n - side) // 2
    z = np.ascontiguousarray(valley["bed"][r0:r0 + side, r0:r0 + side])
    eta = float(z.min()) + 10.0
    out = {}
    for name, datum in (("shifted", "auto"), ("unshifted", None)):
        el = rbs.datum.shift_for_datum(datum, z)
        h0 = np.maximum((eta - el.z_ref) - el.bed.astype(np.float64), 0.0).astype(np.float32)
        st = rbs.State.from_bed(el.bed, 5.0, depth=h0, manning=0.035)
        for _ in range(1000):
            rbs.hl.step(st, rbs.hl.compute_dt(st, 0.45, 30.0))
        u, v = st.velocities_numpy()
        out[name] = (float(max(np.abs(u).max(), np.abs(v).max())), el.z_ref)
        st.close()
    print(f"\n[valley lake at rest 1024^2, 1000 steps] shifted (z_ref {out['shifted'][1]:g} m): "
          f"{out['shifted'][0]:.2e} m/s; unshifted: {out['unshifted'][0]:.2e} m/s "
          f"(target 1e-6, reference gate 1e-4)")
    assert out["shifted"][1] > 0.0
    assert out["shifted"][0] < 1e-4


def test_valley_crop_against_the_oracle(rbs, orc, valley):
    """A 256^2 crop at the head of the valley: the channel at base-flow normal depth,
    the triangular hydrograph (its first 20 minutes, scaled to the crop's share of the
    channel) entering on row 0, the fixed stage at the outlet row, HLLC for 200 steps:
    bitwise the default oracle, within 1e-5 of the libm oracle."""
    n, side = 8192, 256
    c0 = n // 2 - side // 2
    z = np.ascontiguousarray(valley["bed"][:side, c0:c0 + side])
    cols = [j - c0 for j in valley["channel_cols"]]
    q_base = 50.0 / 40.0
    h_n = (q_base * 0.035 / math.sqrt(5e-4)) ** 0.6
    axis = z[:, [side // 2]]
    depth = np.maximum(axis + h_n - z, 0.0).astype(np.float32)
    stage = float(z[-1, side // 2] + h_n)
    hydro = [(0.0, 50.0), (3600.0, 800.0), (14400.0, 50.0)]
    srcs = [((0, j), [(t, q / len(cols)) for t, q in hydro]) for j in cols]
    g = rbs.State.from_bed(z, 5.0, depth=depth, manning=0.035)
    o = orc.State(z, 5.0, depth=depth, manning=0.035)
    o_libm = orc.State(z, 5.0, depth=depth, manning=0.035)
    for s in (g, o, o_libm):
        s.set_open_boundaries({"south": "fixed_stage"}, {"south": [(0.0, stage)]})
    inj = rbs.inflow.InflowInjector([rbs.inflow.Inflow(c, hg) for c, hg in srcs], g)
    comp, comp_l = np.zeros(len(srcs), np.float32), np.zeros(len(srcs), np.float32)
    rows, cc = [0] * len(srcs), [c[1] for c, _ in srcs]
    t = 0.0
    for _ in range(200):
        dt = rbs.hl.compute_dt(g, 0.45, 30.0)
        assert dt == orc.hllc_compute_dt(o, 0.45, 30.0)
        inj.apply(g, t, dt)
        add = np.array([orc.discharge_at(hg, t + 0.5 * dt) * dt / 25.0 for _, hg in srcs],
                       np.float32)
        orc.inject_point_sources_c(o, rows, cc, add, comp)
        orc.inject_point_sources_c(o_libm, rows, cc, add, comp_l)
        rbs.hl.step(g, dt, t=t)
        orc.hllc_step(o, dt, t=t)
        with orc.pow_mode("libm"):
            orc.hllc_step(o_libm, dt, t=t)
        t += dt
    d = [float(np.abs(g.h.numpy() - o.h).max()), float(np.abs(g.hu.numpy() - o.hu).max()),
         float(np.abs(g.hv.numpy() - o.hv).max())]
    dl = [float(np.abs(g.h.numpy() - o_libm.h).max()), float(np.abs(g.hu.numpy() - o_libm.hu).max()),
          float(np.abs(g.hv.numpy() - o_libm.hv).max())]
    wet = int((g.h.numpy() >= H_DRY).sum())
    print(f"\n[valley crop 256^2, {t:.0f} s] default oracle {d}, libm oracle {dl}, wet {wet}")
    assert d == [0.0, 0.0, 0.0]
    assert max(dl) <= 1e-5
    assert wet > 2000
