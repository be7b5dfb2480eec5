"""Sub-grid channels in the oracle and the host geometry code, pinned by the
reference's own tests (solver/core/test_channels.py, validation/test_subgrid_channel.py,
solver/core/test_sediment.py:402-471) and by closed forms: the storage curve's
branches, the column-depth CFL bound, a single two-component face restated in
float64, Manning normal depth of the channel's own section, and exact mass closure.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from rbs.core.channels import (
    ChannelError,
    eta_from_h,
    h_from_eta,
    validate_geometry,
)
from rbs.core.grid import Grid
from rbs.core.sediment import capacity_from_flow

G = 9.81


# --- the storage curve (host, float64) -------------------------------------------
def test_storage_curve_is_continuous_monotone_and_invertible():
    """test_channels.py:42-64."""
    dx, w, d, z = 100.0, 20.0, 3.0, 12.0
    h_bf = w * d / dx
    assert eta_from_h(h_bf, z, w, d, dx) == pytest.approx(z, abs=1e-12)
    assert eta_from_h(h_bf - 1e-9, z, w, d, dx) == pytest.approx(z, abs=1e-6)
    assert eta_from_h(h_bf + 1e-9, z, w, d, dx) == pytest.approx(z, abs=1e-6)
    assert eta_from_h(h_bf / 2, z, w, d, dx) == pytest.approx(z - d / 2)
    assert eta_from_h(h_bf + 0.4, z, w, d, dx) == pytest.approx(z + 0.4)
    h = np.linspace(0.0, 5.0, 501)
    eta = eta_from_h(h, z, w, d, dx)
    assert np.all(np.diff(eta) > 0.0)
    np.testing.assert_allclose(h_from_eta(eta, z, w, d, dx), h, atol=1e-12)
    assert eta_from_h(0.0, z, w, d, dx) == pytest.approx(z - d)


def test_no_channel_is_exactly_the_plain_relation():
    h = np.linspace(0.0, 3.0, 64)
    z = np.full_like(h, 7.25)
    zero = np.zeros_like(h)
    np.testing.assert_array_equal(eta_from_h(h, z, zero, zero, 30.0), z + h)


def test_oracle_storage_curve_and_column_match_the_host_reference(orc):
    """test_channels.py:74-94, plus column_depth = eta - (z - d) on channel cells."""
    rng = np.random.default_rng(7)
    dx = 100.0
    h = rng.uniform(0.0, 2.0, 64).astype(np.float32)
    z = rng.uniform(5.0, 15.0, 64).astype(np.float32)
    w = np.where(rng.random(64) < 0.5, 0.0, rng.uniform(5.0, 40.0, 64)).astype(np.float32)
    d = np.where(w > 0, rng.uniform(0.5, 4.0, 64), 0.0).astype(np.float32)
    got = np.array([orc.eta_subgrid(*a, dx) for a in zip(h, z, w, d)])
    np.testing.assert_allclose(got, eta_from_h(h, z, w, d, dx), rtol=1e-5, atol=1e-5)
    col = np.array([orc.column_depth(a, b, c, dx) for a, b, c in zip(h, w, d)])
    want = np.where(w > 0, eta_from_h(h, z, w, d, dx) - (z.astype(np.float64) - d), h)
    np.testing.assert_allclose(col, want, rtol=1e-5, atol=1e-5)
    # no channel: the plain relation bit for bit
    assert orc.eta_subgrid(0.37, 4.0, 0.0, 0.0, dx) == np.float32(np.float32(0.37) + np.float32(4.0))


# --- geometry validation (host) ----------------------------------------------------
def test_geometry_that_cannot_mean_anything_is_refused():
    """test_channels.py:100-127."""
    grid = Grid(ny=3, nx=4, dx=50.0)
    ok_w = np.full((3, 4), 10.0, np.float32)
    ok_d = np.full((3, 4), 2.0, np.float32)
    ok_n = np.full((3, 4), 0.03, np.float32)
    wide = ok_w.copy()
    wide[1, 1] = 60.0
    with pytest.raises(ChannelError, match="wider than the cell"):
        validate_geometry(wide, ok_d, ok_n, grid)
    half = ok_d.copy()
    half[0, 0] = 0.0
    with pytest.raises(ChannelError, match="width without a bank-full depth"):
        validate_geometry(ok_w, half, ok_n, grid)
    neg = ok_w.copy()
    neg[2, 3] = -1.0
    with pytest.raises(ChannelError, match=">= 0"):
        validate_geometry(neg, ok_d, ok_n, grid)
    with pytest.raises(ChannelError, match="shape"):
        validate_geometry(np.zeros((2, 2), np.float32), ok_d, ok_n, grid)
    bad_n = ok_n.copy()
    bad_n[0, 0] = 0.0
    with pytest.raises(ChannelError, match="manning > 0"):
        validate_geometry(ok_w, ok_d, bad_n, grid)
    nan = ok_d.copy()
    nan[1, 2] = np.nan
    with pytest.raises(ChannelError, match="non-finite"):
        validate_geometry(ok_w, nan, ok_n, grid)


def test_a_hairline_channel_warns_because_it_will_crawl():
    grid = Grid(ny=2, nx=2, dx=200.0)
    with pytest.warns(UserWarning, match="timestep"):
        validate_geometry(np.full((2, 2), 0.01, np.float32), np.ones((2, 2), np.float32),
                          np.full((2, 2), 0.03, np.float32), grid)


def test_no_channel_normalises_to_zero_in_both_fields():
    grid = Grid(ny=2, nx=2, dx=50.0)
    w = np.array([[0.0, 10.0], [0.0, 0.0]], np.float32)
    d = np.array([[0.0, 2.0], [0.0, 0.0]], np.float32)
    w_out, d_out, _ = validate_geometry(w, d, np.full((2, 2), 0.03, np.float32), grid)
    assert w_out[0, 0] == 0.0 and d_out[0, 0] == 0.0
    assert w_out[0, 1] == 10.0 and d_out[0, 1] == 2.0


def test_a_width_of_exactly_dx_is_accepted():
    """A channel filling its cell is still sub-grid (the check is a strict w > dx)."""
    grid = Grid(ny=1, nx=2, dx=50.0)
    w, _, _ = validate_geometry(np.full((1, 2), 50.0), np.ones((1, 2)), np.full((1, 2), 0.03),
                                grid)
    assert w.dtype == np.float32 and w.max() == 50.0


# --- one face, restated in float64 -------------------------------------------------
def _pow43(x):
    return x ** (4.0 / 3.0)


def test_manning_denominator_radius_is_the_section_form(orc):
    """friction.py:35-57: D = 1 + g dt n^2 |q| / (h R^(4/3)); R = h is the wide form."""
    rng = np.random.default_rng(4)
    for _ in range(200):
        q, h, n, dt = rng.uniform(-3, 3), rng.uniform(0.01, 5), rng.uniform(0.01, 0.1), 0.7
        r = h * rng.uniform(0.3, 1.0)
        want = 1.0 + G * dt * n * n * abs(q) / (h * _pow43(r))
        assert orc.manning_denominator_radius(q, h, r, n, 9.81, dt) == pytest.approx(want, rel=2e-6)
        wide = 1.0 + G * dt * n * n * abs(q) / h ** (7.0 / 3.0)
        assert orc.manning_denominator_radius(q, h, h, n, 9.81, dt) == pytest.approx(wide, rel=2e-6)


def _two_cells(orc, *, z, h, w, d, n_fp=0.05, n_ch=0.03, dx=50.0):
    st = orc.State(np.array([z], np.float32), dx, depth=np.array([h], np.float32), manning=n_fp)
    st.arm_channels(np.array([w], np.float32), np.array([d], np.float32),
                    np.full((1, 2), n_ch, np.float32))
    return st


def test_an_in_bank_face_is_the_channel_component_times_its_width_fraction(orc):
    """Both cells below bank full: the floodplain component is dry and the total is
    q_ch * w/dx, q_ch from the channel's own depth and radius A/P (local_inertial.py
    :286-381)."""
    dx, dt = 50.0, 0.5
    st = _two_cells(orc, z=[10.0, 9.95], h=[0.5, 0.3], w=[20.0, 20.0], d=[2.0, 2.0], dx=dx)
    st.qx_ch[0, 1] = 0.3
    st.qx_fp[0, 1] = 0.7  # stale floodplain flow: a dry floodplain face is 0
    orc.li_step(st, dt, limit=False)
    eta = [10.0 - 2.0 + 0.5 * dx / 20.0, 9.95 - 2.0 + 0.3 * dx / 20.0]
    h_ch = min(max(eta), 10.0) - 8.0
    radius = 20.0 * h_ch / (20.0 + 2.0 * h_ch)
    slope = (eta[1] - eta[0]) / dx
    q_ch = (0.3 - G * h_ch * dt * slope) / (1.0 + G * dt * 0.03**2 * 0.3 / (h_ch * _pow43(radius)))
    assert float(st.qx_fp[0, 1]) == 0.0
    assert float(st.qx_ch[0, 1]) == pytest.approx(q_ch, rel=1e-5)
    assert float(st.qx[0, 1]) == pytest.approx(q_ch * 20.0 / dx, rel=1e-5)
    # the narrower section conveys, the higher invert is the sill
    st2 = _two_cells(orc, z=[10.0, 9.95], h=[0.5, 0.3], w=[20.0, 10.0], d=[2.0, 1.5], dx=dx)
    orc.li_step(st2, dt, limit=False)
    eta2 = [8.0 + 0.5 * dx / 20.0, 9.95 - 1.5 + 0.3 * dx / 10.0]
    h_ch2 = min(max(eta2), 10.0) - max(8.0, 8.45)
    r2 = 10.0 * h_ch2 / (10.0 + 2.0 * h_ch2)
    q2 = (0.0 - G * h_ch2 * dt * (eta2[1] - eta2[0]) / dx)  # q_prev = 0: D = 1
    assert float(st2.qx_ch[0, 1]) == pytest.approx(q2, rel=1e-5)
    assert float(st2.qx[0, 1]) == pytest.approx(q2 * 10.0 / dx, rel=1e-5)
    assert r2 < h_ch2  # the wetted perimeter includes the banks


def test_an_overbank_face_mixes_both_components_by_area(orc):
    """Surface above the bank: the floodplain carries eta_max - z_bank and the channel
    saturates at bank full (channels.py:115-131)."""
    dx, dt = 50.0, 0.25
    st = _two_cells(orc, z=[10.0, 9.9], h=[1.5, 1.2], w=[15.0, 15.0], d=[1.0, 1.0], dx=dx)
    st.qx_ch[0, 1] = 0.8
    st.qx_fp[0, 1] = 0.2
    orc.li_step(st, dt, limit=False)
    h_bf = 15.0 * 1.0 / dx
    eta = [10.0 + (1.5 - h_bf), 9.9 + (1.2 - h_bf)]
    slope = (eta[1] - eta[0]) / dx
    h_fp = max(eta) - 10.0
    q_fp = (0.2 - G * h_fp * dt * slope) / (1.0 + G * dt * 0.05**2 * 0.2 / h_fp ** (7.0 / 3.0))
    h_ch = 10.0 - 9.0  # saturated at bank full: min(eta_max, z_bank) - z_ch
    r = 15.0 * h_ch / (15.0 + 2.0 * h_ch)
    q_ch = (0.8 - G * h_ch * dt * slope) / (1.0 + G * dt * 0.03**2 * 0.8 / (h_ch * _pow43(r)))
    frac = 15.0 / dx
    assert float(st.qx_fp[0, 1]) == pytest.approx(q_fp, rel=1e-5)
    assert float(st.qx_ch[0, 1]) == pytest.approx(q_ch, rel=1e-5)
    assert float(st.qx[0, 1]) == pytest.approx(q_fp * (1 - frac) + q_ch * frac, rel=1e-5)


def test_the_limiter_scales_both_components_by_the_totals_donor(orc):
    """limit_qx_channels (local_inertial.py:432-458): a drained donor scales the total
    and both components by the same factor, keeping q = q_fp (1-f) + q_ch f."""
    dx, dt = 10.0, 5.0
    kw = dict(z=[10.0, 9.0], h=[0.05, 0.0], w=[5.0, 5.0], d=[1.0, 1.0], n_ch=0.01, dx=dx)
    st = _two_cells(orc, **kw)
    st.qx_ch[0, 1] = 5.0
    orc.li_step(st, dt, limit=False)
    unlimited = (float(st.qx[0, 1]), float(st.qx_ch[0, 1]), float(st.qx_fp[0, 1]))
    st = _two_cells(orc, **kw)
    st.qx_ch[0, 1] = 5.0
    orc.li_step(st, dt, limit=True)
    b = float(st.beta[0, 0])
    assert 0.0 < b < 1.0
    assert float(st.qx[0, 1]) == np.float32(np.float32(unlimited[0]) * np.float32(b))
    assert float(st.qx_ch[0, 1]) == np.float32(np.float32(unlimited[1]) * np.float32(b))
    assert float(st.h[0, 0]) >= 0.0 and float(st.h[0, 0]) < 1e-6  # drained, not negative


# --- the seam: zero width, the timestep, closed-box mass -----------------------------
def _slope_bed(ny, nx, dx, slope, base=5.0):
    xx = np.broadcast_to(np.arange(nx), (ny, nx))
    return ((nx - 1 - xx) * dx * slope + base).astype(np.float32)


def test_zero_width_channels_reproduce_the_plain_scheme_bitwise(orc):
    """test_channels.py:147-170."""
    ny, nx, dx = 8, 24, 40.0
    bed = _slope_bed(ny, nx, dx, 0.004)
    plain = orc.State(bed, dx, depth=0.05, manning=0.03)
    armed = orc.State(bed, dx, depth=0.05, manning=0.03)
    armed.arm_channels(np.zeros((ny, nx)), np.zeros((ny, nx)))
    for _ in range(60):
        dt_p = orc.li_compute_dt(plain, 0.7, 5.0)
        dt_a = orc.li_compute_dt(armed, 0.7, 5.0)
        assert dt_p == dt_a
        orc.li_step(plain, dt_p, rain=1e-5)
        orc.li_step(armed, dt_a, rain=1e-5)
    np.testing.assert_array_equal(armed.h, plain.h)
    np.testing.assert_array_equal(armed.qx, plain.qx)
    np.testing.assert_array_equal(armed.qy, plain.qy)
    np.testing.assert_array_equal(armed.qx_fp, plain.qx)  # the total is the floodplain's
    assert not armed.qx_ch.any()


def test_timestep_is_set_by_the_water_column_not_the_cell_mean(orc):
    """test_channels.py:173-186."""
    ny, nx, dx = 4, 8, 200.0
    w, d, h0 = 20.0, 3.0, 0.05
    bed = np.full((ny, nx), 10.0, np.float32)
    plain = orc.State(bed, dx, depth=h0)
    armed = orc.State(bed, dx, depth=h0)
    armed.arm_channels(np.full((ny, nx), w), np.full((ny, nx), d))
    dt_plain = orc.li_compute_dt(plain, 0.7, 1e6)
    dt_armed = orc.li_compute_dt(armed, 0.7, 1e6)
    assert dt_plain == pytest.approx(0.7 * dx / math.sqrt(G * h0), rel=1e-5)
    assert dt_armed == pytest.approx(0.7 * dx / math.sqrt(G * h0 * dx / w), rel=1e-5)
    assert dt_armed < dt_plain / 3.0
    # over bank full the column is d + (h - h_bf)
    over = orc.State(bed, dx, depth=0.5)
    over.arm_channels(np.full((ny, nx), w), np.full((ny, nx), d))
    col = d + (0.5 - w * d / dx)
    assert orc.li_compute_dt(over, 0.7, 1e6) == pytest.approx(0.7 * dx / math.sqrt(G * col),
                                                              rel=1e-5)


def test_a_closed_channelled_box_conserves_mass_exactly(orc):
    """test_channels.py:189-209."""
    ny, nx, dx = 10, 20, 50.0
    yy, xx = np.mgrid[0:ny, 0:nx]
    bed = (20.0 + 0.02 * ((yy - ny / 2) ** 2 + (xx - nx / 2) ** 2)).astype(np.float32)
    st = orc.State(bed, dx, depth=0.2, manning=0.03)
    w = np.zeros((ny, nx), np.float32)
    w[ny // 2, :] = 15.0
    d = np.zeros((ny, nx), np.float32)
    d[ny // 2, :] = 2.0
    st.arm_channels(w, d)
    v0 = float(st.h.astype(np.float64).sum())
    for _ in range(300):
        orc.li_step(st, orc.li_compute_dt(st, 0.7, 5.0))
    v1 = float(st.h.astype(np.float64).sum())
    assert np.isfinite(st.h).all()
    assert st.h.min() >= -1e-9
    assert abs(v1 - v0) / v0 < 1e-6
    assert np.abs(st.qx_ch).max() > 0.0  # the channel carried water


# --- the physics gates (validation/test_subgrid_channel.py) --------------------------
def normal_depth(discharge, width, slope, manning):
    """Manning normal depth of a rectangular section (bisection on Q(h), R = A/P)."""
    def q_of(h):
        area = width * h
        return area * (area / (width + 2.0 * h)) ** (2.0 / 3.0) * math.sqrt(slope) / manning

    lo, hi = 1e-6, 100.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if q_of(mid) < discharge else (lo, mid)
    return 0.5 * (lo + hi)


_EAST_OPEN = {"east": "open", "west": "closed", "north": "closed", "south": "closed"}


def _run_steady(orc, st, sources, *, steps, dt_max):
    """validation/test_subgrid_channel.py:79-98 with the reference's point-source
    injection (inflow.py:91-200): midpoint discharge, Kahan-added depth."""
    led = orc.MassLedger.from_state(st)
    rows = [s[0][0] for s in sources]
    cols = [s[0][1] for s in sources]
    comp = np.zeros(len(sources), np.float32)
    area = st.cell_area
    t = 0.0
    for _ in range(steps):
        dt = orc.li_compute_dt(st, 0.7, dt_max)
        add = np.array([orc.discharge_at(hg, t + 0.5 * dt) * dt / area for _, hg in sources],
                       np.float32)
        led.add_inflow(float(add.astype(np.float64).sum()) * area)
        orc.inject_point_sources_c(st, rows, cols, add, comp)
        orc.li_step(st, dt)
        t += dt
    led.record(st, t)
    return led, t


def test_subgrid_channel_reaches_manning_normal_depth(orc):
    """validation/test_subgrid_channel.py:101-140: within 5% of the section's normal
    depth, in bank, mass inside the gate."""
    nx, dx = 100, 50.0
    slope, n_ch, width, bankfull, q = 0.001, 0.03, 20.0, 4.0, 30.0
    st = orc.State(_slope_bed(1, nx, dx, slope), dx, depth=0.0, manning=0.05)
    st.arm_channels(np.full((1, nx), width), np.full((1, nx), bankfull), np.full((1, nx), n_ch))
    st.set_open_boundaries(_EAST_OPEN)
    led, _ = _run_steady(orc, st, [((0, 0), [(0.0, q), (1.0e9, q)])], steps=3000, dt_max=10.0)
    h_n = normal_depth(q, width, slope, n_ch)
    column = st.h[0] * dx / width
    median = float(np.median(column[10:nx - 10]))
    assert np.isfinite(column).all()
    assert led.max_rel_error < orc.MASS_GATE
    assert median < bankfull
    assert median == pytest.approx(h_n, rel=0.05)


def test_overbank_spill_leaves_the_channel_and_conserves_mass(orc):
    """validation/test_subgrid_channel.py:143-180."""
    ny, nx, dx = 7, 30, 50.0
    st = orc.State(_slope_bed(ny, nx, dx, 0.0005), dx, depth=0.0, manning=0.05)
    row = ny // 2
    w = np.zeros((ny, nx), np.float32)
    d = np.zeros((ny, nx), np.float32)
    w[row, :] = 20.0
    d[row, :] = 1.0
    st.arm_channels(w, d, np.full((ny, nx), 0.03, np.float32))
    led, _ = _run_steady(orc, st, [((row, 2), [(0.0, 60.0), (1.0e9, 60.0)])], steps=1200,
                         dt_max=5.0)
    h_bf = 20.0 * 1.0 / dx
    floodplain = np.delete(st.h, row, axis=0)
    assert np.isfinite(st.h).all()
    assert st.h.min() >= -1e-9
    assert st.h[row].max() > h_bf
    assert (floodplain > 1e-3).sum() > 0
    assert led.max_rel_error < orc.MASS_GATE


def test_coarse_subgrid_channel_matches_the_resolved_channel(orc):
    """validation/test_subgrid_channel.py:183-255: the reach claim, 100x fewer cells."""
    length, slope, n_ch, width, bankfull, q = 2000.0, 0.0008, 0.035, 40.0, 2.0, 40.0
    dxf = 10.0
    nyf, nxf = 10, int(length / dxf)
    bed_f = _slope_bed(nyf, nxf, dxf, slope)
    rows = slice(3, 7)
    bed_f[rows, :] -= bankfull
    fine = orc.State(bed_f, dxf, depth=0.0, manning=n_ch)
    fine.set_open_boundaries(_EAST_OPEN)
    per = q / 4.0
    fine_led, _ = _run_steady(orc, fine, [((r, 0), [(0.0, per), (1.0e9, per)]) for r in range(3, 7)],
                              steps=6000, dt_max=5.0)
    dxc = 100.0
    nxc = int(length / dxc)
    coarse = orc.State(_slope_bed(1, nxc, dxc, slope), dxc, depth=0.0, manning=n_ch)
    coarse.arm_channels(np.full((1, nxc), width), np.full((1, nxc), bankfull),
                        np.full((1, nxc), n_ch))
    coarse.set_open_boundaries(_EAST_OPEN)
    coarse_led, _ = _run_steady(orc, coarse, [((0, 0), [(0.0, q), (1.0e9, q)])], steps=2000,
                                dt_max=10.0)

    def interior(a):
        k = len(a)
        return a[k // 5:-k // 5]

    fine_depth = float(np.median(interior(fine.h[rows].mean(axis=0))))
    coarse_depth = float(np.median(interior(coarse.h[0] * dxc / width)))
    fine_vol = float(fine.h.astype(np.float64).sum()) * dxf * dxf
    coarse_vol = float(coarse.h.astype(np.float64).sum()) * dxc * dxc
    h_wide = (q / width * n_ch / math.sqrt(slope)) ** 0.6
    h_section = normal_depth(q, width, slope, n_ch)
    assert fine_led.max_rel_error < orc.MASS_GATE
    assert coarse_led.max_rel_error < orc.MASS_GATE
    assert fine_depth == pytest.approx(h_wide, rel=0.10)
    assert coarse_depth == pytest.approx(h_section, rel=0.10)
    assert coarse_depth == pytest.approx(fine_depth, rel=0.12)
    assert coarse_vol == pytest.approx(fine_vol, rel=0.20)


# --- channel bedload (solver/core/test_sediment.py:402-471) ---------------------------
def _sed_state(orc, ny, nx, dx, z, eta, w, d, n_fp, n_ch, d50):
    st = orc.State(z, dx, depth=0.0, manning=n_fp)
    st.arm_channels(np.full((ny, nx), w, np.float32), np.full((ny, nx), d, np.float32),
                    np.full((ny, nx), n_ch, np.float32))
    st.arm_sediment(d50)
    st.eta[...] = eta
    return st


def test_a_channel_face_carries_its_conveyance_fraction(orc):
    """test_sediment.py:402-443: surface below the bank, so the floodplain component is
    dry and q_s dx == q_s,ch w, with the channel's own depth and radius A/P."""
    ny, nx, dx, dt = 1, 4, 200.0, 1.0
    w, bank, n_ch, d50 = 20.0, 1.0, 0.028, 0.008
    z = np.full((ny, nx), 10.0, np.float32)
    h_ch = 0.5
    st = _sed_state(orc, ny, nx, dx, z, z - bank + h_ch, w, bank, 0.05, n_ch, d50)
    st.qx_ch[...] = 1.8
    orc.accumulate_transport(st, dt)
    radius = w * h_ch / (w + 2.0 * h_ch)
    qs_channel = float(capacity_from_flow(1.8, h_ch, n_ch, d50, radius=radius))
    got = float(st.qs_int_x[0, 1])
    assert qs_channel > 0.0
    assert got * dx == pytest.approx(qs_channel * w, rel=1e-4)
    naive = float(capacity_from_flow(1.8, h_ch * w / dx, n_ch, d50))
    assert naive > 10.0 * qs_channel


def test_the_channel_transport_collapses_onto_the_plain_one_at_zero_width(orc):
    """test_sediment.py:446-471, bit for bit."""
    ny, nx, dx, dt = 3, 6, 50.0, 3.0
    z = (np.full((ny, nx), 10.0) - 0.05 * np.arange(nx)[None, :]).astype(np.float32)
    rng = np.random.default_rng(3)
    eta = z + rng.uniform(0.3, 1.5, (ny, nx)).astype(np.float32)
    qx = rng.uniform(-2.0, 2.0, (ny, nx + 1)).astype(np.float32)
    plain = orc.State(z, dx, depth=0.0, manning=0.033)
    plain.arm_sediment(0.006)
    plain.eta[...] = eta
    plain.qx[...] = qx
    orc.accumulate_transport(plain, dt)
    chan = _sed_state(orc, ny, nx, dx, z, eta, 0.0, 0.0, 0.033, 0.0, 0.006)
    chan.qx_fp[...] = qx
    orc.accumulate_transport(chan, dt)
    assert np.array_equal(plain.qs_int_x, chan.qs_int_x)
    assert np.array_equal(plain.qs_int_y, chan.qs_int_y)


def test_channel_step_with_bedload_leaves_the_water_unchanged(orc):
    """sediment.py:444-461: the integral reads the post-limiter components and writes
    nothing the step goes on to use."""
    ny, nx, dx = 9, 40, 25.0
    bed = _slope_bed(ny, nx, dx, 0.004)
    w = np.zeros((ny, nx), np.float32)
    d = np.zeros((ny, nx), np.float32)
    w[4, :] = 10.0
    d[4, :] = 1.2
    a = orc.State(bed, dx, depth=0.1, manning=0.04)
    b = orc.State(bed, dx, depth=0.1, manning=0.04)
    for s in (a, b):
        s.arm_channels(w, d, np.full((ny, nx), 0.03, np.float32))
    b.arm_sediment(0.002)
    for _ in range(80):
        dt = orc.li_compute_dt(a, 0.7, 5.0)
        orc.li_step(a, dt, rain=2e-5)
        orc.li_step(b, dt, rain=2e-5)
    np.testing.assert_array_equal(a.h, b.h)
    np.testing.assert_array_equal(a.qx_ch, b.qx_ch)
    assert np.abs(b.qs_int_x).max() > 0.0
