"""Pin the oracle's HLLC restatement (hllc.py) before trusting it as the checker.

The published wet-bed Stoker figures pin the whole step (test_oracle_li.py).
Here: properties the Riemann solver must have whatever its code says --
consistency F(U, U) = F(U), mirror antisymmetry, supersonic upwinding, exact zero
mass flux at a reflective wall, dry-dry zero -- and the scheme-level invariants of
the reference's validation suite (lake at rest on a bumpy bed, fixed-stage
equilibrium, banked open-edge mass).
"""

from __future__ import annotations

import math

import numpy as np
import pytest

G = 9.81


def phys(h, u, v):
    return h * u, h * u * u + 0.5 * G * h * h, h * u * v


def test_flux_consistency_at_uniform_state(orc):
    """validation/test_hllc1d.py:92-98."""
    for h, u, v in [(2.0, 1.5, -0.7), (0.3, -0.2, 0.4), (1.0, 0.0, 0.0), (5.0, 3.0, 1.0)]:
        got = orc.hllc_flux(h, u, v, h, u, v)
        want = phys(h, u, v)
        np.testing.assert_allclose(got, want, rtol=2e-6, atol=1e-7)


def test_transverse_momentum_rides_the_contact_upwind(orc):
    """validation/test_hllc1d.py:101-105."""
    fh, _, fv = orc.hllc_flux(1.0, 3.0, 5.0, 1.0, 3.0, -5.0)
    assert fv == pytest.approx(fh * 5.0, rel=1e-6)
    fh, _, fv = orc.hllc_flux(1.0, -3.0, 5.0, 1.0, -3.0, -5.0)
    assert fv == pytest.approx(fh * -5.0, rel=1e-6)


def test_dry_over_dry_gives_zero_flux(orc):
    """validation/test_hllc1d.py:86-89 (dry = h <= 1e-3)."""
    assert orc.hllc_flux(0.0, 0.0, 0.0, 0.0, 0.0, 0.0) == (0.0, 0.0, 0.0)
    assert orc.hllc_flux(5e-4, 1.0, 0.0, 1e-3, -1.0, 0.0) == (0.0, 0.0, 0.0)


def test_mirror_antisymmetry(orc):
    """Reflecting x -> -x swaps the states and negates normal velocities: the mass
    flux changes sign and the normal-momentum flux is unchanged."""
    rng = np.random.default_rng(1)
    for _ in range(200):
        hl, hr = rng.uniform(0.01, 3.0, 2)
        ul, ur, vl, vr = rng.uniform(-3, 3, 4)
        a = orc.hllc_flux(hl, ul, vl, hr, ur, vr)
        b = orc.hllc_flux(hr, -ur, vr, hl, -ul, vl)
        assert a[0] == pytest.approx(-b[0], rel=1e-5, abs=1e-6)
        assert a[1] == pytest.approx(b[1], rel=1e-5, abs=1e-6)


def test_supersonic_flow_takes_the_upwind_physical_flux(orc):
    """sL >= 0 (Fr > 1 to the right): F = F(U_L) exactly (hllc.py:237-239)."""
    h, u = 0.5, 10.0  # c = 2.2 m/s
    got = orc.hllc_flux(h, u, 0.3, 0.4, 9.0, -1.0)
    assert got == pytest.approx(tuple(np.float32(x) for x in phys(np.float32(h), np.float32(u),
                                                                     np.float32(0.3))), rel=1e-6)


def test_reflective_wall_passes_no_mass(orc):
    """The wall ghost (h, -u) | (h, u) has exactly zero mass flux (hllc.py:453-459)."""
    for h, u in [(1.0, 0.7), (0.2, -1.3), (3.0, 2.5)]:
        fh, fm, ft = orc.hllc_flux(h, -u, 0.4, h, u, 0.4)
        assert fh == 0.0 and ft == 0.0
        assert fm > 0.5 * G * h * h * 0.999 if u < 0 else fm > 0.0


def test_hll_flux_between_the_wave_speeds(orc):
    """Subsonic wet/wet: the mass flux lies between the two one-sided fluxes' bracket
    and reproduces the HLL formula with the two-rarefaction speeds (Toro 10.48)."""
    hl, ul, hr, ur = 1.0, 0.5, 0.6, 0.2
    cl, cr = math.sqrt(G * hl), math.sqrt(G * hr)
    hs = (0.5 * (cl + cr) + 0.25 * (ul - ur)) ** 2 / G
    us = 0.5 * (ul + ur) + cl - cr
    sl = min(ul - cl, us - math.sqrt(G * hs))
    sr = max(ur + cr, us + math.sqrt(G * hs))
    fh = (sr * hl * ul - sl * hr * ur + sl * sr * (hr - hl)) / (sr - sl)
    got = orc.hllc_flux(hl, ul, 0.0, hr, ur, 0.0)
    assert got[0] == pytest.approx(fh, rel=1e-5)


def _bumpy_bed(ny, nx, dx):
    yy, xx = np.mgrid[0:ny, 0:nx]
    return (0.5 * np.sin(xx * 0.2) + 0.002 * xx * dx + 0.3 * np.cos(yy * 0.3)).astype(np.float32)


def test_lake_at_rest_stays_flat(orc):
    """validation/test_hllc_2d.py:36-55: max|u,v| < 1e-4, surface drift < 1e-3."""
    z = _bumpy_bed(30, 40, 10.0)
    eta0 = float(z.max()) + 2.0
    st = orc.State(z, 10.0, depth=np.maximum(eta0 - z, 0.0), manning=0.03)
    for _ in range(200):
        orc.hllc_step(st, orc.hllc_compute_dt(st, 0.45, 5.0))
    u = np.where(st.h >= 1e-3, st.hu / np.maximum(st.h, 1e-3), 0)
    v = np.where(st.h >= 1e-3, st.hv / np.maximum(st.h, 1e-3), 0)
    assert max(np.abs(u).max(), np.abs(v).max()) < 1e-4
    assert np.abs((st.h + z) - eta0).max() < 1e-3


def test_fixed_stage_at_equilibrium_is_inert(orc):
    """validation/test_fixed_stage.py:79-111: a stage edge at the lake's level leaves
    it at rest (max|u,v| < 1e-4) and banks < 1e-3 of the volume."""
    ny = nx = 24
    yy, xx = np.mgrid[0:ny, 0:nx].astype(np.float64)
    bed = (0.02 * xx + 0.3 * np.sin(0.6 * xx) * np.cos(0.4 * yy)).astype(np.float32)
    surface = float(bed.max()) + 0.4
    h0 = (surface - bed.astype(np.float64)).astype(np.float32)
    st = orc.State(bed, 10.0, depth=h0, manning=0.0)
    st.set_open_boundaries({"east": "fixed_stage"}, {"east": [(0.0, surface)]})
    led = orc.MassLedger.from_state(st)
    t = 0.0
    while t < 600.0 - 1e-6:
        dt = min(orc.hllc_compute_dt(st, 0.45, 30.0), 600.0 - t)
        orc.hllc_step(st, dt, t=t)
        t += dt
    led.record(st, t)
    u = np.where(st.h >= 1e-3, st.hu / np.maximum(st.h, 1e-3), 0)
    v = np.where(st.h >= 1e-3, st.hv / np.maximum(st.h, 1e-3), 0)
    assert max(np.abs(u).max(), np.abs(v).max()) < 1e-4
    assert led.max_rel_error < orc.MASS_GATE
    assert abs(st.loss_volume(100.0)) < 1e-3 * led.series[0]["volume"]


@pytest.mark.parametrize("edge", ["north", "south", "east", "west"])
def test_stage_edge_fills_a_dry_domain_from_each_edge(orc, edge):
    """validation/test_fixed_stage.py:114-147: water enters and settles at the stage."""
    ny = nx = 20
    if edge == "east":
        bed = np.linspace(1.0, 0.0, nx)[None, :].repeat(ny, axis=0)
    elif edge == "west":
        bed = np.linspace(0.0, 1.0, nx)[None, :].repeat(ny, axis=0)
    elif edge == "south":
        bed = np.linspace(1.0, 0.0, ny)[:, None].repeat(nx, axis=1)
    else:
        bed = np.linspace(0.0, 1.0, ny)[:, None].repeat(nx, axis=1)
    bed = bed.astype(np.float32)
    st = orc.State(bed, 10.0, depth=0.0, manning=0.03)
    st.set_open_boundaries({edge: "fixed_stage"}, {edge: [(0.0, 0.5)]})
    led = orc.MassLedger.from_state(st)
    t = 0.0
    while t < 1800.0 - 1e-6:
        dt = min(orc.hllc_compute_dt(st, 0.45, 2.0), 1800.0 - t)
        orc.hllc_step(st, dt, t=t)
        t += dt
    led.record(st, t)
    eta = st.h + bed
    wet = st.h > 10 * 1e-3
    assert led.max_rel_error < orc.MASS_GATE
    assert -st.loss_volume(100.0) > 0.0
    assert wet.sum() == pytest.approx((bed < 0.5 - 0.05).sum(), rel=0.25)
    assert float(np.abs(eta[wet] - 0.5).max()) < 0.05


def test_stage_above_the_surface_fills_and_the_ledger_closes(orc):
    """A stage above the interior drives inflow, banked as negative loss."""
    z = np.zeros((6, 20), np.float32)
    st = orc.State(z, 5.0, depth=0.5, manning=0.03)
    st.set_open_boundaries({"west": "fixed_stage"}, {"west": [(0.0, 0.5), (100.0, 1.0)]})
    led = orc.MassLedger.from_state(st)
    t = 0.0
    for _ in range(100):
        dt = orc.hllc_compute_dt(st, 0.45, 5.0)
        orc.hllc_step(st, dt, t=t)
        t += dt
    led.record(st, t)
    assert st.loss_cum.sum() < 0.0  # water came in
    assert led.max_rel_error < 1e-6


def test_open_edge_drain_banks_the_outflow(orc):
    """Transmissive edges + banking keep the ledger closed on a draining tilt."""
    x = np.arange(30)[None, :] * np.ones((8, 1))
    st = orc.State((-0.05 * x).astype(np.float32), 2.0, depth=0.3, manning=0.03)
    st.set_open_boundaries({"east": "open", "west": "open"})
    led = orc.MassLedger.from_state(st)
    for _ in range(150):
        orc.hllc_step(st, orc.hllc_compute_dt(st, 0.45, 5.0))
    led.record(st, 1.0)
    assert st.loss_cum.sum() > 0.0
    assert led.max_rel_error < 1e-6


def test_hllc_determinism(orc):
    """validation/test_hllc_2d.py:58-75: bitwise rerun (with rain)."""
    z = _bumpy_bed(16, 20, 10.0)
    yy, xx = np.mgrid[0:16, 0:20]
    d0 = (0.5 + 0.3 * np.sin(xx * 0.5 + yy * 0.3)).astype(np.float32)

    def run():
        st = orc.State(z, 10.0, depth=d0, manning=0.03)
        for _ in range(60):
            orc.hllc_step(st, orc.hllc_compute_dt(st, 0.45, 5.0), rain=1e-5)
        return st.h.copy(), st.hu.copy()

    a, b = run(), run()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_dry_bed_ritter_front(orc):
    """validation/test_dam_break.py:142-175 diagnostic: mass gate hard, shape < 0.5."""
    nx, dx, t_end = 800, 0.5, 6.0
    x = (np.arange(nx) + 0.5 - nx / 2.0) * dx
    st = orc.State(np.zeros((1, nx)), dx, depth=np.where(x < 0, 1.0, 1e-3).reshape(1, nx),
                   manning=0.02)
    led = orc.MassLedger.from_state(st)
    t = 0.0
    while t < t_end - 1e-9:
        dt = min(orc.hllc_compute_dt(st, 0.45, t_end), t_end - t)
        orc.hllc_step(st, dt)
        t += dt
    led.record(st, t)
    h_ref, _ = orc.ritter(x, t_end, 1.0)
    inter = np.abs(x) < 0.9 * 200
    nrmse = math.sqrt(np.mean((st.h[0][inter] - h_ref[inter]) ** 2))
    assert led.max_rel_error < 1e-6
    assert nrmse < 0.05


# --- the reference's own float64 1-D solver (validation/hllc1d.py) ---------------------
_VEC = None


def _vectors():
    global _VEC
    if _VEC is None:
        from pathlib import Path

        _VEC = np.load(Path(__file__).parent / "golden" / "hllc1d_vectors.npz")
    return _VEC


def test_flux_matches_the_reference_1d_flux(orc):
    """hllc.py:191-256 against hllc1d.py:46-115 on 400 face states, dry sides included
    (tests/golden/hllc1d_vectors.npz, written by make_hllc1d_vectors.py): the float32
    flux within 5e-6 of the float64 one, relative to 1 + |F|."""
    d = _vectors()
    args = [d[f"flux/{k}"] for k in ("hL", "uL", "vL", "hR", "uR", "vR")]
    got = np.array([orc.hllc_flux(*(float(a[i]) for a in args)) for i in range(args[0].size)])
    for j, k in enumerate(("Fh", "Fhu", "Fhv")):
        ref = d[f"flux/{k}"]
        assert (np.abs(got[:, j] - ref) / (1.0 + np.abs(ref))).max() < 5e-6, k


def run_1d_case(step, compute_dt, make_state, d, name):
    """One of the golden wet cases on a 1 x N row with transmissive ends, stepping to
    t_end with dt = min(CFL dt, t_end - t) as solve_1d does (hllc1d.py:197-251)."""
    g = lambda k: d[f"{name}/{k}"]  # noqa: E731
    nx = g("h0").size
    st = make_state(g("z").reshape(1, nx).astype(np.float32), float(g("dx")),
                    g("h0").reshape(1, nx).astype(np.float32),
                    (g("h0") * g("u0")).reshape(1, nx).astype(np.float32), float(g("manning_n")))
    t, t_end = 0.0, float(g("t_end"))
    while t < t_end - 1e-9:
        dt = min(compute_dt(st), t_end - t)
        step(st, dt, t)
        t += dt
    return st, g("h"), g("u")


@pytest.mark.parametrize("name", ["flow_over_bump", "wet_dam_break_bumpy_slope",
                                  "lake_over_two_bumps"])
def test_wet_flows_match_the_reference_1d_solver(orc, name):
    """The 2-D float32 HLLC step on one row against the reference's float64 1-D solver:
    flow over a bump, a wet dam break on a bumpy slope, a lake over two bumps; depth
    within 5e-6 m and velocity within 3e-5 m/s."""

    def make(z, dx, h0, hu0, n):
        st = orc.State(z, dx, depth=h0, manning=n)
        st.arm_hllc()
        st.hu[...] = hu0
        st.set_open_boundaries({"west": "open", "east": "open"})
        return st

    st, h_ref, u_ref = run_1d_case(lambda s, dt, t: orc.hllc_step(s, dt, t=t),
                                   lambda s: orc.hllc_compute_dt(s, 0.45, 1e9), make, _vectors(), name)
    h = st.h[0].astype(np.float64)
    u = np.where(h > 1e-3, st.hu[0] / np.maximum(h, 1e-3), 0.0)
    assert h.min() > 0.5  # wet everywhere: the regime both operators share
    assert np.abs(h - h_ref).max() < 5e-6
    assert np.abs(u - u_ref).max() < 3e-5
