"""HLLC + Exner in the oracle: the coupling the library adds (the reference runs
morphology on the local-inertial scheme only, config.py:451-461), pinned by the
invariants the reference's morphology gates rest on rather than by a reference run:

* below the Shields threshold the bed change is bit-exact zero, and just above it the
  same reach moves (validation/test_morphology_gates.py:85-147);
* a bump migrates at the analytical bed-wave celerity at the flow the run delivers,
  ``bed_celerity / (1 - Fr^2)`` for the full shallow-water equations, by the
  reference's shape estimator and gate width (validation/bedwave.py,
  validation/test_bed_wave.py:159-211);
* the sediment ledger closes: interior fluxes telescope, so the bed change plus what
  the bounds banked sums to zero (massbalance.py SedimentLedger);
* a lake at rest stays at rest with morphology armed (the transport writes nothing
  the step reads), moves no bed, and stays at rest after a uniform bed shift: the
  well-balanced reconstruction holds on the moved bed.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, replace

import numpy as np
import pytest

from rbs.core.sediment import SHIELDS_CRITICAL, bed_celerity, capacity_from_flow, shields_from_flow
from rbs.scheduler import fill_span

EPS_T = 1e-6


@dataclass(frozen=True)
class Reach:
    """The reference's bed-wave design point (bedwave.py:280-320): one row, +x flow."""

    dx: float = 2.5
    nx: int = 240
    slope: float = 0.002
    manning: float = 0.035
    d50: float = 0.008
    porosity: float = 0.4
    q: float = 2.5
    amp: float = 0.015
    sigma_m: float = 15.0
    bump_cell: int = 80
    migration_cells: float = 16.0
    interval_s: float = 45.0
    warmup_s: float = 1200.0
    outlet_bed: float = 5.0

    @property
    def depth(self) -> float:  # wide-channel Manning normal depth
        return (self.q * self.manning / math.sqrt(self.slope)) ** 0.6

    @property
    def shields(self) -> float:
        return float(shields_from_flow(self.q, self.depth, self.manning, self.d50))

    def at_shields(self, ratio: float) -> "Reach":
        return replace(self, d50=self.d50 * self.shields / (ratio * SHIELDS_CRITICAL))

    def celerity_at(self, q: float, h: float) -> float:
        return float(bed_celerity(q, h, self.manning, self.d50, self.porosity))

    @property
    def activations(self) -> int:
        run = self.migration_cells * self.dx / self.celerity_at(self.q, self.depth)
        return max(1, int(round(run / self.interval_s)))

    def plane(self) -> np.ndarray:
        j = np.arange(self.nx, dtype=np.float64)
        return self.outlet_bed + (self.nx - 1 - j) * self.dx * self.slope

    def bump(self) -> np.ndarray:
        j = np.arange(self.nx, dtype=np.float64)
        return self.amp * np.exp(-0.5 * ((j - self.bump_cell) / (self.sigma_m / self.dx)) ** 2)

    def bed(self) -> np.ndarray:
        return (self.plane() + self.bump()).reshape(1, self.nx).astype(np.float32)

    @property
    def window(self) -> tuple[int, int]:
        pad = int(math.ceil(3.0 * self.sigma_m / self.dx))
        return self.bump_cell - pad, min(self.bump_cell + int(math.ceil(self.migration_cells)) + pad,
                                         self.nx)


def xcorr_lag(ref, final, lo, hi, max_lag) -> float:
    """Downstream shift (cells) of ``ref`` that best matches ``final`` in the window,
    refined by a parabola through the peak (bedwave.py:707-754)."""
    max_lag = min(max_lag, lo)
    seg = final[lo:hi]
    corr = np.array([float(np.dot(ref[lo - k:hi - k], seg)) for k in range(max_lag + 1)])
    k = int(np.argmax(corr))
    assert k < max_lag, "the bump moved past the search window"
    if k == 0:
        return 0.0
    y0, y1, y2 = corr[k - 1], corr[k], corr[k + 1]
    den = y0 - 2.0 * y1 + y2
    return float(k) if den == 0.0 else float(k) + 0.5 * (y0 - y2) / den


def drive(orc, fx: Reach, *, n_act: int, morphology: bool = True):
    """The reference's fixture driver (bedwave.py:615-705) on the oracle's HLLC step:
    constant inflow at the head, the toe at normal depth, water-only warm-up, then
    sediment armed and an Exner activation every interval with the end cells pinned."""
    orc.set_threads(1)
    st = orc.State(fx.bed(), fx.dx, depth=fx.depth, manning=fx.manning)
    # the toe held at normal depth: HLLC's open edge is transmissive, which does not
    # control a subcritical reach shorter than its backwater length h/S
    st.set_open_boundaries({"east": "fixed_stage"},
                           {"east": [(0.0, fx.outlet_bed + fx.depth)]})
    led = orc.MassLedger.from_state(st)
    comp = np.zeros(1, np.float32)
    area = st.cell_area
    lo = np.full((1, fx.nx), -np.inf, np.float32)
    hi = np.full((1, fx.nx), np.inf, np.float32)
    lo[0, [0, -1]] = hi[0, [0, -1]] = 0.0
    dz = np.zeros((1, fx.nx), np.float64)
    un = np.zeros((1, fx.nx), np.float64)
    z0 = None
    end = fx.warmup_s + n_act * fx.interval_s
    t, acts = 0.0, 0
    while t < end - EPS_T:
        dt = orc.hllc_compute_dt(st, 0.45, 5.0)
        edge = fx.warmup_s if t < fx.warmup_s - EPS_T else fx.warmup_s + (acts + 1) * fx.interval_s
        if not morphology:  # no activations to land on (bedwave.py:658-666)
            edge = end
        dt = fill_span(min(edge, end) - t, dt)
        if morphology and z0 is None and t >= fx.warmup_s - EPS_T:
            st.arm_sediment(fx.d50)
            z0 = st.z.copy()
        add = np.array([fx.q * fx.dx * dt / area], np.float32)
        led.add_inflow(float(add.astype(np.float64)[0]) * area)
        orc.inject_point_sources_c(st, [0], [0], add, comp)
        orc.hllc_step(st, dt, t=t)
        t += dt
        if z0 is not None and t >= fx.warmup_s + (acts + 1) * fx.interval_s - EPS_T:
            orc.exner_update(st.qs_int_x, st.qs_int_y, lo, hi, fx.dx, fx.porosity, dz, un)
            orc.rebuild_bed(z0, dz, st.z)
            st.qs_int_x[...] = 0.0
            st.qs_int_y[...] = 0.0
            acts += 1
    led.record(st, t)
    return st, led, dz, un, acts


@pytest.fixture(scope="module")
def design_run(orc):
    fx = Reach()
    return fx, drive(orc, fx, n_act=fx.activations)


def test_the_bump_migrates_at_the_analytical_celerity(design_run):
    fx, (st, led, dz, un, acts) = design_run
    assert acts == fx.activations
    # Water mass is the HLLC scheme's float32 round-off, not the coupling (Exner moves
    # no water): this reach passes 16x its storage through in 30 000 steps and closes
    # to 8.4e-6 with morphology off, 1.2e-6 with it on.
    assert led.max_rel_error < 1e-5
    interior = slice(30, fx.nx - 30)
    h = float(np.median(st.h[0, interior]))
    q = float(np.median(st.hu[0, interior]))
    assert h == pytest.approx(fx.depth, rel=0.02) and q == pytest.approx(fx.q, rel=0.02)
    c_flow = fx.celerity_at(q, h)  # rigid-lid celerity (sediment.py:747-801)
    # With advection the surface over a short bump dips by dh = -dz / (1 - Fr^2)
    # (steady shallow water, friction negligible over sigma << h/S), so the full
    # equations' bed wave runs 1/(1 - Fr^2) faster than the rigid-lid value the
    # local-inertial scheme (no advection) reproduces (bedwave.py:73-76).
    froude = q / h / math.sqrt(9.81 * h)
    c_swe = c_flow / (1.0 - froude ** 2)
    final = np.asarray(st.z[0], np.float64) - fx.plane()
    lo, hi = fx.window
    lag = xcorr_lag(fx.bump(), final, lo, hi, int(math.ceil(2 * fx.migration_cells)))
    celerity = lag * fx.dx / (acts * fx.interval_s)
    crest = float(final[lo:hi].max())
    print(f"\n[hllc bed wave] c = {celerity:.3e} m/s: {celerity / c_swe:.3f} of c_b/(1-Fr^2) "
          f"(Fr {froude:.3f}), {celerity / c_flow:.3f} of rigid-lid c_b; crest kept "
          f"{crest / fx.amp:.2f}; mass {led.max_rel_error:.1e}")
    assert celerity == pytest.approx(c_swe, rel=0.20)  # the gate width of test_bed_wave.py:209
    assert celerity > 1.1 * c_flow  # and it is the advective response, not the rigid lid
    assert 0.5 < crest / fx.amp < 1.05


def test_the_sediment_ledger_closes(design_run):
    """Interior faces telescope: sum(dz_cum) + sum(banked) = 0 to round-off."""
    fx, (st, led, dz, un, acts) = design_run
    gross = float(np.abs(dz).sum())
    assert gross > 1e-3
    assert abs(float(dz.sum()) + float(un.sum())) / gross < 1e-11
    assert dz[0, 0] == 0.0 and dz[0, -1] == 0.0  # the pinned ends


def test_just_under_the_threshold_the_bed_change_is_bit_exact_zero(orc):
    fx = Reach().at_shields(0.90)
    st, led, dz, un, acts = drive(orc, fx, n_act=20)
    assert acts == 20 and led.max_rel_error < 1e-5
    assert (dz == 0.0).all() and (un == 0.0).all()
    assert np.array_equal(st.z, fx.bed())
    assert not st.qs_comp_x.any()


def test_just_over_the_threshold_the_same_reach_moves(orc):
    fx = Reach().at_shields(1.20)
    st, led, dz, un, acts = drive(orc, fx, n_act=20)
    assert led.max_rel_error < 1e-5
    assert int((dz != 0.0).sum()) >= fx.nx - 4 and np.abs(dz).max() > 1e-3


def _lake(orc, ny=24, nx=40, seed=3):
    rng = np.random.default_rng(seed)
    z = (2.0 * rng.uniform(0, 1, (ny, nx))).astype(np.float32)
    st = orc.State(z, 5.0, depth=np.maximum(np.float32(3.0) - z, 0).astype(np.float32),
                   manning=0.03)
    return st


def test_a_lake_at_rest_moves_no_bed_and_is_untouched_by_the_transport(orc):
    plain, armed = _lake(orc), _lake(orc)
    armed.arm_sediment(np.full(armed.z.shape, 1e-4, np.float32))
    for _ in range(200):
        dt = orc.hllc_compute_dt(plain, 0.45, 1.0)
        assert dt == orc.hllc_compute_dt(armed, 0.45, 1.0)
        orc.hllc_step(plain, dt)
        orc.hllc_step(armed, dt)
    assert np.array_equal(plain.h, armed.h) and np.array_equal(plain.hu, armed.hu)
    assert not armed.qs_int_x.any() and not armed.qs_int_y.any()
    assert np.abs(armed.hu).max() / 1.0 < 1e-3


def test_a_uniform_bed_shift_keeps_the_lake_at_rest(orc):
    """Bounds lo = hi = c make every cell's cumulative change exactly c: the bed and
    the surface shift together and the well-balanced reconstruction holds."""
    st = _lake(orc)
    st.arm_sediment(np.full(st.z.shape, 1e-4, np.float32))
    z0 = st.z.copy()
    for _ in range(50):
        orc.hllc_step(st, orc.hllc_compute_dt(st, 0.45, 1.0))
    before = float(np.abs(st.hu).max())
    shift = np.full(st.z.shape, 0.25, np.float32)
    dz = np.zeros(st.z.shape, np.float64)
    un = np.zeros(st.z.shape, np.float64)
    orc.exner_update(st.qs_int_x, st.qs_int_y, shift, shift, 5.0, 0.4, dz, un)
    orc.rebuild_bed(z0, dz, st.z)
    assert np.all(dz == np.float64(np.float32(0.25)))
    for _ in range(200):
        orc.hllc_step(st, orc.hllc_compute_dt(st, 0.45, 1.0))
    after = float(np.abs(st.hu).max())
    wet = st.h > 1e-3
    eta = (st.h + st.z)[wet]
    assert after < 1e-3 and after <= max(10 * before, 1e-5)
    assert eta.max() - eta.min() < 1e-4


def test_capacity_reference_is_the_law_the_kernels_take(orc):
    """The float64 host law and the oracle's float32 face capacity agree."""
    rng = np.random.default_rng(5)
    for _ in range(100):
        q, h, n, d = rng.uniform(0.5, 4), rng.uniform(0.3, 3), rng.uniform(0.02, 0.05), rng.uniform(5e-4, 0.01)
        want = float(capacity_from_flow(q, h, n, d))
        got = orc.face_capacity(q, h, h, n, d)
        assert got == pytest.approx(want, rel=2e-5, abs=1e-12)
