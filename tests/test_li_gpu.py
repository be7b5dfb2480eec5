"""Local-inertial HIP path (librbs through the C-ABI) against the oracle.

Tolerance (north star): after 1000 fixed-dt steps max|dh| <= 1e-5 m and
max|dq| <= 1e-5 m^2/s.  The fused kernel follows the reference's float32
operation order with no FMA contraction and evaluates wp.pow with the reference's
float32 exponents as the correctly rounded power, by the same operation sequence as
the oracle's default mode, so against that mode the fields are equal bit for bit.
Against the oracle's libm mode (glibc powf, what the reference's Warp-CPU runs
evaluated) they differ only where glibc misses the last bit, within the tolerance.
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
    from rbs.core import local_inertial, massbalance, state

    _lib.load()
    assert _lib.device_count() >= 1
    return type("R", (), dict(lib=_lib, li=local_inertial, State=state.State,
                              MassLedger=massbalance.MassLedger))


def _pair(rbs, orc, bed, dx, depth, manning, *, threads=8):
    orc.set_threads(threads)
    g = rbs.State.from_bed(bed, dx, depth=depth, manning=manning)
    o = orc.State(bed, dx, depth=depth, manning=manning)
    return g, o


def _dt(rbs, orc, g, o, alpha, dt_max):
    """The step both sides take: the device's CFL dt (states may differ by ulps, so
    the oracle's own dt is only checked to agree closely)."""
    dt = rbs.li.compute_dt(g, alpha, dt_max)
    do = orc.li_compute_dt(o, alpha, dt_max)
    assert abs(dt - do) <= 1e-6 * dt, (dt, do)
    return dt


def _cmp(g, o, tol=TOL):
    dh = np.abs(g.h.numpy() - o.h).max()
    dqx = np.abs(g.qx.numpy() - o.qx).max()
    dqy = np.abs(g.qy.numpy() - o.qy).max()
    assert dh <= tol, f"max|dh| {dh:.3e}"
    assert dqx <= tol and dqy <= tol, f"max|dq| {max(dqx, dqy):.3e}"
    return dh, max(dqx, dqy)


def _dambreak_512(rbs, orc):
    ny = nx = 512
    depth = np.where(np.arange(nx)[None, :] < 256, 1.0, 0.1) * np.ones((ny, 1))
    g, o = _pair(rbs, orc, np.zeros((ny, nx)), 1.0, depth, 0.01)
    o_libm = orc.State(np.zeros((ny, nx)), 1.0, depth=depth, manning=0.01)
    return g, o, o_libm


def _stoker_score(orc, row, t):
    """validation/test_dam_break.py:93-133's scores on the 512^2 dam break's centreline row."""
    nx = row.size
    x = (np.arange(nx) + 0.5 - 256.0) * 1.0
    h_ref, _ = orc.stoker(x, t, 1.0, 0.1)
    inter = np.abs(x) < 0.9 * 256
    lvl = 0.5 * (0.1 + orc.stoker_star_depth(1.0, 0.1))
    fref = orc.stoker_front_position(t, 1.0, 0.1)
    h = row.astype(np.float64)
    nrmse = math.sqrt(np.mean((h[inter] - h_ref[inter]) ** 2)) / 0.9
    return nrmse, abs(x[h > lvl].max() - fref) / fref


def test_dambreak_512_1000_steps_parity_mass_and_stoker(rbs, orc):
    """512x512 dam break (h=1 for x<256, 0.1 else), n=0.01, closed, the fixed dt
    alpha dx / sqrt(g h_max(t=0)) for 1000 steps.  The wave reflections raise h_max to
    1.65 m, above which this dt exceeds the CFL bound (about step 100 on), so the run
    is sensitive to every last bit: the HIP path equals the default oracle bit for bit,
    and the libm oracle (the reference's glibc powf) as well -- no power in this run
    lands on one of glibc's misrounded inputs."""
    g, o, o_libm = _dambreak_512(rbs, orc)
    dt = 0.7 * 1.0 / math.sqrt(9.81 * 1.0)  # compute_dt at t = 0 (h_max = 1)
    assert rbs.li.compute_dt(g, 0.7, 30.0) == dt == orc.li_compute_dt(o, 0.7, 30.0)
    led = rbs.MassLedger.from_state(g)
    for k in range(250):
        rbs.li.step(g, dt)
    orc.li_run_fixed(o, dt, 250)
    assert _cmp(g, o, tol=0.0) == (0.0, 0.0)
    # Stoker along the centreline row at t = 250 dt (both waves still interior).  The
    # bore h_R/h_L = 0.1 is far stronger than the reference gate's 0.3 (the next test
    # runs that gate); the LI scheme, which drops advection, scores 0.19 here.
    nrmse, front_err = _stoker_score(orc, g.h.numpy()[256], 250 * dt)
    assert nrmse < 0.25 and front_err < 0.15
    for k in range(750):
        rbs.li.step(g, dt)
    orc.li_run_fixed(o, dt, 750)
    dh, dq = _cmp(g, o, tol=0.0)
    with orc.pow_mode("libm"):
        orc.li_run_fixed(o_libm, dt, 1000)
    dh_l, dq_l = _cmp(g, o_libm)
    rec = led.record(g, 1000 * dt)
    assert rec.rel_error < 1e-7, f"mass closure {rec.rel_error:.2e}"
    print(f"\n[dam break 512^2 parity] default oracle max|dh|={dh:.2e} max|dq|={dq:.2e}; libm oracle "
          f"{dh_l:.2e} / {dq_l:.2e}; mass={rec.rel_error:.2e} stoker nRMSE={nrmse:.4f} "
          f"front={front_err:.4f}")


def test_dambreak_512_at_a_stable_dt(rbs, orc):
    """The 512^2 dam break at a fixed dt that keeps the CFL bound for all 1000 steps: alpha dx /
    sqrt(g h_peak) with h_peak = 2 m above the 1.65 m the reflections reach.  Parity with
    both oracle modes; the Stoker scores at the same time as the fixed-dt run, reported."""
    g, o, o_libm = _dambreak_512(rbs, orc)
    dt = 0.7 * 1.0 / math.sqrt(9.81 * 2.0)
    k_stoker = int(round(250 * 0.7 / math.sqrt(9.81) / dt))  # t = 55.9 s
    for k in range(k_stoker):
        rbs.li.step(g, dt)
    nrmse, front_err = _stoker_score(orc, g.h.numpy()[256], k_stoker * dt)
    for k in range(1000 - k_stoker):
        rbs.li.step(g, dt)
    orc.li_run_fixed(o, dt, 1000)
    with orc.pow_mode("libm"):
        orc.li_run_fixed(o_libm, dt, 1000)
    dh, dq = _cmp(g, o, tol=0.0)
    dh_l, dq_l = _cmp(g, o_libm)
    assert float(g.h.numpy().max()) < 2.0  # the dt stayed within the CFL bound
    # the LI scheme drops advection and lags this strong bore (h_R/h_L = 0.1): 0.187 and
    # 0.180 here; the reference's band (0.10, 0.15) is for its 0.3 bore, passed above
    assert nrmse < 0.25 and front_err < 0.25
    print(f"\n[dam break 512^2, stable dt {dt:.4f} s] default oracle {dh:.2e} / {dq:.2e}; libm oracle "
          f"{dh_l:.2e} / {dq_l:.2e}; stoker nRMSE={nrmse:.4f} front={front_err:.4f}")


@pytest.fixture(scope="module")
def basin4096():
    """The rain-on-terrain basin at its full 4096^2 (seed 20261203)."""
    from rbs import terrain

    return terrain.rain_on_terrain(4096, seed=20261203)


def test_rain_basin_crop_512_1000_fixed_dt_steps(rbs, orc, basin4096):
    """A 512^2 crop of the rain-on-terrain basin's generators (its centre), all edges
    open, rain field with Kahan debt, infiltration, initially dry, 1000 steps of 1 s:
    bitwise the default oracle (h, qx, qy and the Kahan debt) and within 1e-5 of the
    libm oracle."""
    sl = np.s_[1792:2304, 1792:2304]
    T = {k: (v[sl].copy() if isinstance(v, np.ndarray) else v) for k, v in basin4096.items()}
    g, o = _pair(rbs, orc, T["bed"], T["dx"], 0.0, T["manning"])
    o_libm = orc.State(T["bed"], T["dx"], depth=0.0, manning=T["manning"])
    for s in (g, o, o_libm):
        s.set_rain_field(T["rain"])
        s.set_infiltration(T["infil"])
        s.set_open_boundaries({e: "open" for e in ("north", "south", "east", "west")})
        s.arm_source_compensation()
    for _ in range(1000):
        rbs.li.step(g, 1.0, rain_scale=1.0)
    orc.li_run_fixed(o, 1.0, 1000, rain_scale=1.0)
    with orc.pow_mode("libm"):
        orc.li_run_fixed(o_libm, 1.0, 1000, rain_scale=1.0)
    dh, dq = _cmp(g, o, tol=0.0)
    assert np.array_equal(g.h_comp.numpy(), o.h_comp)
    dh_l, dq_l = _cmp(g, o_libm)
    wet = int((g.h.numpy() >= 1e-3).sum())
    assert wet > 50_000
    print(f"\n[rain basin 512^2 crop] default oracle {dh:.2e} / {dq:.2e}; libm oracle {dh_l:.2e} / "
          f"{dq_l:.2e}; wet cells {wet}")


def test_rain_basin_4096_device_driven(rbs, basin4096):
    """The whole 4096^2 basin, device-driven (the headline kernel, as the bench runs it)
    from dry through 1800 s of storm, several hundred steps: mass closes within 1e-7 of the
    throughput (inflow + outflow), every depth is finite and non-negative."""
    from rbs.driver import DeviceRun
    from rbs.scheduler import MultiRateScheduler

    T = basin4096
    st = rbs.State.from_bed(T["bed"], T["dx"], depth=0.0, manning=T["manning"])
    st.set_rain_field(T["rain"])
    st.set_infiltration(T["infil"])
    st.set_open_boundaries({e: "open" for e in ("north", "south", "east", "west")})
    st.arm_source_compensation()
    rain_sum = float(T["rain"].astype(np.float64).sum())
    sched = MultiRateScheduler(end_time=1800.0, output_every=600.0)
    run = DeviceRun(st, "local_inertial", sched, alpha=0.7, dt_max=30.0, rain_is_field=True,
                    rain_duration=7200.0, rain_field_sum=rain_sum)
    steps = sum(run.run_until(seg.t_stop) for seg in sched.segments())
    led = st.ledger()
    h = st.h.numpy()
    assert steps >= 200
    assert np.isfinite(h).all() and h.min() >= -1e-6
    assert led.t == pytest.approx(1800.0, abs=1e-9)
    # the device ledger's rain is the field's volume over the clock, in float64
    assert led.rain_in == pytest.approx(rain_sum * T["dx"] ** 2 * 1800.0, rel=1e-9)
    inflow, outflow = led.rain_in, led.infil_out + led.edge_out
    closure = abs(inflow - outflow - led.volume) / (inflow + outflow)
    assert closure <= 1e-7, closure
    print(f"\n[rain basin 4096^2] {steps} steps to 1800 s, mass closure {closure:.2e} of throughput, "
          f"h_max {h.max():.3f} m")


def test_li_wet_bed_stoker_gate_on_gpu(rbs, orc):
    """validation/test_dam_break.py:93-133 for LI: nRMSE < 0.10, front < 0.15, mass < 1e-6;
    the published run gives 0.074 / 9.5 % (M1-water-moves.md:231-233)."""
    nx, dx, t_end = 800, 0.5, 8.0
    x = (np.arange(nx) + 0.5 - nx / 2.0) * dx
    g = rbs.State.from_bed(np.zeros((1, nx)), dx, depth=np.where(x < 0, 1.0, 0.3).reshape(1, nx),
                           manning=0.01)
    led = rbs.MassLedger.from_state(g)
    t = 0.0
    while t < t_end - 1e-9:
        dt = min(rbs.li.compute_dt(g, 0.7, t_end), t_end - t)
        rbs.li.step(g, dt)
        t += dt
    rec = led.record(g, t)
    h = g.h.numpy()[0].astype(np.float64)
    h_ref, _ = orc.stoker(x, t_end, 1.0, 0.3)
    inter = np.abs(x) < 0.9 * 200
    nrmse = math.sqrt(np.mean((h[inter] - h_ref[inter]) ** 2)) / 0.7
    lvl = 0.5 * (0.3 + orc.stoker_star_depth(1.0, 0.3))
    fr = orc.stoker_front_position(t_end, 1.0, 0.3)
    front = abs(x[h > lvl].max() - fr) / fr
    assert nrmse < 0.10 and front < 0.15 and rec.rel_error < 1e-6
    assert round(nrmse, 3) == 0.074 and round(front, 3) == 0.095


@pytest.mark.parametrize("shape", [(1, 1), (1, 37), (41, 1), (2, 2), (67, 133), (130, 65)])
def test_odd_shapes_match(rbs, orc, shape):
    """Tiles that overhang the grid, single rows/columns, 1x1."""
    rng = np.random.default_rng(sum(shape))
    ny, nx = shape
    z = rng.uniform(0, 1, shape)
    h = rng.uniform(0.0, 1.0, shape)
    n = rng.uniform(0.02, 0.08, shape)
    g, o = _pair(rbs, orc, z, 2.0, h, n)
    for _ in range(40):
        dt = _dt(rbs, orc, g, o, 0.7, 1.0)
        rbs.li.step(g, dt)
        orc.li_step(o, dt)
    _cmp(g, o, tol=1e-6)


@pytest.mark.parametrize("limit", [True, False])
def test_friction_ratio_extremes_match_the_oracle(rbs, orc, limit):
    """The face's friction quotient g dt n^2 |q| / h^float32(7/3) is the IEEE division
    without its operand scaling, which acts only when the quotient vanishes in 1 + c/p
    or |q| passes 1e19 m^2/s.  One step over discharges of 0, -0, subnormal and tiny
    magnitudes up to 1e15 m^2/s, depths from 1e-3 m (the dry threshold, so every face
    is wet) to 1e3 m and Manning n from 0 to 0.2, bitwise the oracle (whose division
    is C's)."""
    rng = np.random.default_rng(21)
    ny, nx = 64, 96
    mags = np.float32([0.0, 1e-45, 3e-42, 1e-38, 2e-30, 1e-20, 1e-10, 1e-4, 0.5, 30.0,
                       1e3, 1e6, 1e9, 1e12, 1e15])
    depths = np.float32([1e-3, 1.001e-3, 2e-3, 1e-2, 0.1, 1.0, 10.0, 100.0, 1e3])

    def q_field(shape):
        return (rng.choice(mags, shape) * rng.choice(np.float32([1.0, -1.0]), shape)).astype(
            np.float32)

    h = rng.choice(depths, (ny, nx)).astype(np.float32)
    z = np.zeros((ny, nx), np.float32)
    n = rng.choice(np.float32([0.0, 0.01, 0.05, 0.2]), (ny, nx)).astype(np.float32)
    qx, qy = q_field((ny, nx + 1)), q_field((ny + 1, nx))
    g, o = _pair(rbs, orc, z, 5.0, h, n)
    g.qx, g.qy = qx, qy
    o.qx[...] = qx
    o.qy[...] = qy
    rbs.li.step(g, 1e-3, limit=limit)
    orc.li_step(o, 1e-3, limit=limit)
    for a, b in ((g.qx.numpy(), o.qx), (g.qy.numpy(), o.qy), (g.h.numpy(), o.h)):
        assert np.array_equal(a.view(np.uint32), np.asarray(b, np.float32).view(np.uint32))


def test_sources_sinks_open_edges_parity(rbs, orc):
    """Rain field (Kahan), infiltration, all edges open, on a sloped random bed."""
    rng = np.random.default_rng(5)
    ny, nx = 96, 160
    yy, xx = np.mgrid[0:ny, 0:nx]
    z = (0.02 * xx + 0.01 * yy + 0.3 * rng.uniform(0, 1, (ny, nx))).astype(np.float32)
    rain = rng.uniform(0.5, 2.0, (ny, nx)).astype(np.float32) * 1e-5
    infil = rng.uniform(0, 10, (ny, nx)).astype(np.float32) / 1000 / 3600
    n = rng.uniform(0.02, 0.1, (ny, nx))
    g, o = _pair(rbs, orc, z, 10.0, 0.0, n)
    for s in (g, o):
        s.set_rain_field(rain)
        s.set_infiltration(infil)
        s.set_open_boundaries({e: "open" for e in ("north", "south", "east", "west")})
        s.arm_source_compensation()
    led_g = rbs.MassLedger.from_state(g)
    led_o = orc.MassLedger.from_state(o)
    rain_sum = float(rain.astype(np.float64).sum())
    t = 0.0
    for k in range(600):
        scale = 1.0 if k < 400 else 0.0
        dt = _dt(rbs, orc, g, o, 0.7, 5.0)
        rbs.li.step(g, dt, rain_scale=scale)
        orc.li_step(o, dt, rain_scale=scale)
        if scale:
            led_g.add_inflow(rain_sum * dt * 100.0)
            led_o.add_inflow(rain_sum * dt * 100.0)
        t += dt
    _cmp(g, o)
    np.testing.assert_allclose(g.h_comp.numpy(), o.h_comp, atol=1e-7)
    rg, ro = led_g.record(g, t), led_o.record(o, t)
    assert rg.outflow_cum == pytest.approx(ro["outflow_cum"], rel=1e-9)
    assert rg.rel_error < 1e-6 and ro["rel_error"] < 1e-6


def test_kahan_sub_ulp_and_canary(rbs):
    """test_sources.py:45-86 through the HIP path."""
    rain, steps = 4.0e-8, 2000
    plain = rbs.State.from_bed(np.zeros((24, 24)), 100.0, depth=1.0, manning=0.03)
    comp = rbs.State.from_bed(np.zeros((24, 24)), 100.0, depth=1.0, manning=0.03)
    comp.arm_source_compensation()
    for _ in range(steps):
        rbs.li.step(plain, 1.0, rain=rain, limit=False)
        rbs.li.step(comp, 1.0, rain=rain, limit=False)
    got_plain = plain.h.numpy().astype(np.float64).mean() - 1.0
    got_comp = comp.h.numpy().astype(np.float64).mean() - 1.0
    assert got_plain == pytest.approx(0.0, abs=1e-9)
    assert got_comp == pytest.approx(rain * steps, rel=2e-3)
    assert np.any(comp.h_comp.numpy() != 0.0)


def test_rain_field_is_read_per_cell(rbs):
    """test_sources.py:179-199."""
    st = rbs.State.from_bed(np.zeros((24, 24)), 100.0, depth=1.0, manning=0.03)
    field = np.zeros((24, 24), np.float32)
    field[:, :12] = 4.0e-8
    st.set_rain_field(field)
    st.arm_source_compensation()
    rbs.li.step(st, 1.0, rain_scale=1.0)
    comp = st.h_comp.numpy()
    assert np.all(comp[:, :12] != 0.0) and np.all(comp[:, 12:] == 0.0)


def test_rest_flux_sign_and_known_dt(rbs):
    """test_local_inertial.py:21-43, 132-139."""
    st = rbs.State.from_bed(np.zeros((8, 8)), 10.0, depth=1.5, manning=0.03)
    for _ in range(20):
        rbs.li.step(st, 1.0)
    assert np.all(st.h.numpy() == np.float32(1.5))
    st2 = rbs.State.from_bed(np.zeros((1, 2)), 1.0, depth=np.array([[2.0, 1.0]]), manning=0.03)
    rbs.li.step(st2, 0.01)
    qx = st2.qx.numpy()
    assert qx[0, 0] == 0.0 and qx[0, 2] == 0.0 and qx[0, 1] > 0.0
    st3 = rbs.State.from_bed(np.zeros((4, 4)), 10.0, depth=4.0)
    assert rbs.li.compute_dt(st3, 0.7, 30.0) == 0.7 * 10.0 / math.sqrt(9.81 * 4.0)


def test_determinism_and_variants_are_bitwise(rbs):
    """Two runs match bit for bit; every tile variant gives the same bits."""
    rng = np.random.default_rng(9)
    z = rng.uniform(0, 2, (200, 300)).astype(np.float32)
    h = rng.uniform(0, 1, (200, 300)).astype(np.float32)
    count = np.zeros(1, np.int32)
    import ctypes

    rbs.lib.call("rbs_li_variant_count", count.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    outs = []
    for v in [0, 0] + list(range(1, int(count[0]))):
        st = rbs.State.from_bed(z, 5.0, depth=h, manning=0.04)
        rbs.lib.call("rbs_set_li_variant", st._handle, v)
        for _ in range(30):
            rbs.li.step(st, rbs.li.compute_dt(st, 0.7, 2.0))
        outs.append((st.h.numpy(), st.qx.numpy(), st.qy.numpy()))
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


def test_variants_with_every_source_and_bedload_are_bitwise(rbs):
    """Rain field (Kahan), infiltration, two open edges and the bedload integral:
    every kernel variant (both families) writes the same bits."""
    import ctypes

    from rbs.core.sediment import arm_sediment

    rng = np.random.default_rng(12)
    ny, nx = 173, 251  # ragged against every tile shape
    yy, xx = np.mgrid[0:ny, 0:nx]
    z = (0.004 * (nx - xx) + 0.3 * rng.uniform(0, 1, (ny, nx))).astype(np.float32)
    h = np.maximum(0.6 - 0.5 * z, 0.0).astype(np.float32)
    rain = rng.uniform(1, 3, (ny, nx)).astype(np.float32) * 1e-5
    infil = rng.uniform(0, 5, (ny, nx)).astype(np.float32) * 1e-6
    d50 = rng.uniform(5e-4, 4e-3, (ny, nx)).astype(np.float32)
    count = np.zeros(1, np.int32)
    rbs.lib.call("rbs_li_variant_count", count.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    outs = []
    for v in range(int(count[0])):
        st = rbs.State.from_bed(z, 5.0, depth=h, manning=0.035)
        st.set_rain_field(rain)
        st.set_infiltration(infil)
        st.set_open_boundaries({"west": "open", "south": "open"})
        st.arm_source_compensation()
        arm_sediment(st, d50, 0.4)
        rbs.lib.call("rbs_set_li_variant", st._handle, v)
        for _ in range(40):
            rbs.li.step(st, rbs.li.compute_dt(st, 0.7, 2.0), rain_scale=1.0)
        sed = st.sediment
        outs.append((st.h.numpy(), st.qx.numpy(), st.qy.numpy(), st.h_comp.numpy(),
                     sed.qs_int_x.numpy(), sed.qs_comp_x.numpy(), sed.qs_int_y.numpy(),
                     sed.qs_comp_y.numpy()))
        st.close()
    names = ["h", "qx", "qy", "h_comp", "qs_int_x", "qs_comp_x", "qs_int_y", "qs_comp_y"]
    for v, o in enumerate(outs[1:], start=1):
        for name, a, b in zip(names, outs[0], o):
            assert np.array_equal(a, b), f"variant {v}: {name} differs"
    assert np.abs(outs[0][4]).max() > 0  # transport happened


def test_device_driven_run_matches_host_loop(rbs):
    """rbs_run_until == the reference's host loop (compute_dt, fill_span, step), bitwise."""
    from rbs.driver import DeviceRun
    from rbs.scheduler import MultiRateScheduler

    rng = np.random.default_rng(2)
    z = rng.uniform(0, 0.5, (128, 160)).astype(np.float32)
    rain_m_s = 20.0 / 1000 / 3600
    sched = MultiRateScheduler(end_time=600.0, output_every=150.0, events=[400.0])

    host = rbs.State.from_bed(z, 10.0, depth=0.05, manning=0.03)
    host.arm_source_compensation()
    dev = rbs.State.from_bed(z, 10.0, depth=0.05, manning=0.03)
    dev.arm_source_compensation()
    led_h = rbs.MassLedger.from_state(host)
    led_d = rbs.MassLedger.from_state(dev)
    n_host = 0
    for tick in sched.ticks(lambda: rbs.li.compute_dt(host, 0.7, 30.0)):
        rain = rain_m_s if tick.t0 < 400.0 - 1e-6 else 0.0
        rbs.li.step(host, tick.dt, rain=rain, t=tick.t0)
        if rain > 0:
            led_h.add_rain_step(rain, tick.dt, host.grid.n_cells)
        n_host += 1
    run = DeviceRun(dev, "local_inertial", sched, alpha=0.7, dt_max=30.0, rain_m_s=rain_m_s,
                    rain_duration=400.0)
    n_dev = 0
    for seg in sched.segments():
        n_dev += run.run_until(seg.t_stop)
    assert n_dev == n_host
    assert np.array_equal(host.h.numpy(), dev.h.numpy())
    assert np.array_equal(host.qx.numpy(), dev.qx.numpy())
    rh, rd = led_h.record(host, 600.0), led_d.record(dev, 600.0)
    assert rd.inflow_cum == pytest.approx(rh.inflow_cum, rel=1e-12)
    assert dev.ledger().t == pytest.approx(600.0, abs=1e-9)
    assert rd.rel_error < 1e-6


def test_row_blocks_at_full_width_equal_the_strips(rbs):
    """The headline kernel family at a width where row blocks end ragged (9 blocks of 480
    columns over 4000, the last with three idle waves) and a height that is not a
    multiple of the block: 30 steps with every source armed, bitwise the strip kernel's
    fields, dt sequence and wave-speed maximum."""
    import ctypes

    from rbs.autotune import variant_table

    table = variant_table()
    rows = [r["variant"] for r in table if r["family"] == "rows"]
    strip = next(r["variant"] for r in table if r["family"] == "strip")
    rng = np.random.default_rng(21)
    ny, nx = 1000, 4000
    yy, xx = np.mgrid[0:ny, 0:nx]
    z = (0.0005 * (nx - xx) + 0.001 * yy + 0.5 * rng.uniform(0, 1, (ny, nx))).astype(np.float32)
    h = np.maximum(1.5 - 0.7 * z, 0.0).astype(np.float32)
    rain = rng.uniform(0, 3, (ny, nx)).astype(np.float32) * 1e-5
    infil = rng.uniform(0, 4, (ny, nx)).astype(np.float32) * 1e-6
    outs = []
    for v in [strip] + rows:
        st = rbs.State.from_bed(z, 10.0, depth=h, manning=0.035)
        st.set_rain_field(rain)
        st.set_infiltration(infil)
        st.set_open_boundaries({"east": "open", "south": "open"})
        st.arm_source_compensation()
        rbs.lib.call("rbs_set_li_variant", st._handle, ctypes.c_int32(v))
        dts = []
        for _ in range(30):
            dt = rbs.li.compute_dt(st, 0.7, 5.0)
            dts.append(dt)
            rbs.li.step(st, dt, rain_scale=1.0)
        outs.append((v, dts, st.h.numpy(), st.qx.numpy(), st.qy.numpy(), st.h_comp.numpy()))
        st.close()
    _, dts0, *f0 = outs[0]
    for v, dts, *f in outs[1:]:
        assert dts == dts0, v
        for name, a, b in zip(("h", "qx", "qy", "h_comp"), f0, f):
            assert np.array_equal(a, b), f"variant {v}: {name} differs"
    assert np.abs(outs[0][3]).max() > 0.0
