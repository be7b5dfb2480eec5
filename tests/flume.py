"""The bed-wave flume of the reference's morphology gates (validation/bedwave.py:280-705),
driven through librbs: one row of 240 cells of 2.5 m on a 0.002 slope, a constant
2.5 m^2/s inflow at the head and an open toe, a 15 mm Gaussian bump 80 cells in,
1200 s of water-only warm-up and then an Exner activation every 45 s with the end cells
pinned.  Test infrastructure (not collected): the local-inertial gates in
tests/test_bed_wave_gpu.py and the clamp-ripple gates use it.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, replace

import numpy as np

from rbs.core.sediment import (SHIELDS_CRITICAL, bed_celerity, capacity_from_flow,
                               morphological_courant, shields_from_flow)
from rbs.scheduler import EPS_T, fill_span

DT_MAX = 5.0  # bedwave.py:588: inert at this scale (the CFL step is about 0.46 s)


@dataclass(frozen=True)
class Flume:
    """Geometry and design flow (bedwave.py:280-306)."""

    dx: float = 2.5
    nx: int = 240
    slope: float = 0.002
    manning: float = 0.035
    d50: float = 0.008
    porosity: float = 0.4
    q: float = 2.5  # m^2/s, per unit width
    amp: float = 0.015
    sigma_m: float = 15.0
    bump_cell: int = 80
    migration_cells: float = 16.0
    interval_s: float = 45.0
    warmup_s: float = 1200.0
    outlet_bed: float = 5.0
    pinned: int = 1

    @property
    def depth(self) -> float:
        """Wide-channel Manning normal depth (q n / sqrt(S))^(3/5)."""
        return (self.q * self.manning / math.sqrt(self.slope)) ** 0.6

    @property
    def froude(self) -> float:
        return self.q / self.depth / math.sqrt(9.81 * self.depth)

    @property
    def shields(self) -> float:
        return float(shields_from_flow(self.q, self.depth, self.manning, self.d50))

    @property
    def capacity(self) -> float:
        return float(capacity_from_flow(self.q, self.depth, self.manning, self.d50))

    def at_shields(self, ratio: float) -> "Flume":
        """The same reach with the grain size that puts it at ratio * theta_c: theta is
        exactly inversely proportional to d50, and d50 does not enter the hydraulics."""
        return replace(self, d50=self.d50 * self.shields / (ratio * SHIELDS_CRITICAL))

    def celerity_at(self, q: float, h: float) -> float:
        return float(bed_celerity(q, h, self.manning, self.d50, self.porosity))

    @property
    def celerity(self) -> float:
        return self.celerity_at(self.q, self.depth)

    @property
    def courant(self) -> float:
        return morphological_courant(self.celerity, self.interval_s, self.dx)

    @property
    def activations(self) -> int:
        """Whole intervals covering the design migration (bedwave.py:404-417)."""
        run = self.migration_cells * self.dx / self.celerity
        return max(1, int(round(run / self.interval_s)))

    def plane(self) -> np.ndarray:
        """The unperturbed bed, float64 (falling towards the open toe)."""
        j = np.arange(self.nx, dtype=np.float64)
        return self.outlet_bed + (self.nx - 1 - j) * self.dx * self.slope

    def bump(self) -> np.ndarray:
        j = np.arange(self.nx, dtype=np.float64)
        return self.amp * np.exp(-0.5 * ((j - self.bump_cell) / (self.sigma_m / self.dx)) ** 2)

    def bed(self) -> np.ndarray:
        return (self.plane() + self.bump()).reshape(1, self.nx).astype(np.float32)

    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        """Free bed change but for the pinned end cells (the flume's sediment BC)."""
        lo = np.full((1, self.nx), -np.inf, np.float32)
        hi = np.full((1, self.nx), np.inf, np.float32)
        k = self.pinned
        lo[0, :k] = hi[0, :k] = lo[0, self.nx - k:] = hi[0, self.nx - k:] = 0.0
        return lo, hi

    @property
    def window(self) -> tuple[int, int]:
        """Cells the estimators look at: three sigma about the bump's whole path."""
        pad = int(math.ceil(3.0 * self.sigma_m / self.dx))
        hi = self.bump_cell + int(math.ceil(self.migration_cells)) + pad
        return self.bump_cell - pad, min(hi, self.nx)

    @property
    def interior(self) -> slice:
        """Cells clear of the head cell and the toe, where the flow is the design flow."""
        pad = max(8, int(math.ceil(2.0 * self.sigma_m / self.dx)))
        return slice(pad, self.nx - pad)


# --- estimators (bedwave.py:712-754) ---------------------------------------------------
def parabolic_peak(a: np.ndarray) -> float:
    """Position (cells) of the largest value, refined by a parabola through it."""
    k = int(np.argmax(a))
    if k in (0, len(a) - 1):
        return float(k)
    y0, y1, y2 = (float(a[k + d]) for d in (-1, 0, 1))
    curv = y0 - 2.0 * y1 + y2
    return float(k) if curv == 0.0 else k + 0.5 * (y0 - y2) / curv


def positive_centroid(a: np.ndarray) -> float:
    w = np.clip(a, 0.0, None)
    s = float(w.sum())
    return 0.0 if s <= 0.0 else float((w * np.arange(len(w))).sum() / s)


def xcorr_lag(ref: np.ndarray, final: np.ndarray, lo: int, hi: int, max_lag: int) -> float:
    """Downstream shift (cells) of ``ref`` that correlates best with ``final`` over the
    window [lo, hi), by integer lag refined with a parabola; refuses a peak at the limit."""
    max_lag = min(max_lag, lo)
    seg = final[lo:hi]
    corr = np.array([float(np.dot(ref[lo - k:hi - k], seg)) for k in range(max_lag + 1)])
    if int(np.argmax(corr)) == max_lag:
        raise ValueError(f"the correlation peaks at its {max_lag}-cell search limit")
    return parabolic_peak(corr)


@dataclass
class Migration:
    """The bump's displacement by three estimators (bedwave.py:233-278)."""

    xcorr_cells: float
    peak_cells: float
    centroid_cells: float
    elapsed_s: float
    dx: float
    crest_m: float
    background_m: float
    amp0_m: float

    def speed(self, cells: float) -> float:
        return cells * self.dx / self.elapsed_s

    @property
    def retained(self) -> float:
        return self.crest_m / self.amp0_m

    @property
    def signal_to_background(self) -> float:
        return self.crest_m / self.background_m if self.background_m > 0 else math.inf


def measure(fx: Flume, z_final: np.ndarray, elapsed_s: float) -> Migration:
    """Each estimator on the initial and the final bed, differenced (bedwave.py:522-553)."""
    lo, hi = fx.window
    ref = fx.bump()
    final = np.asarray(z_final, np.float64).reshape(fx.nx) - fx.plane()
    k = fx.pinned
    outside = np.concatenate([np.abs(final[k:lo]), np.abs(final[hi:fx.nx - k])])
    return Migration(
        xcorr_cells=xcorr_lag(ref, final, lo, hi, int(math.ceil(2.0 * fx.migration_cells))),
        peak_cells=parabolic_peak(final[lo:hi]) - parabolic_peak(ref[lo:hi]),
        centroid_cells=positive_centroid(final[lo:hi]) - positive_centroid(ref[lo:hi]),
        elapsed_s=elapsed_s, dx=fx.dx, crest_m=float(final[lo:hi].max()),
        background_m=float(outside.max()) if outside.size else 0.0, amp0_m=fx.amp)


@dataclass
class FlumeRun:
    """What one run leaves on the host (bedwave.py:594-612)."""

    bed: np.ndarray
    depth: np.ndarray
    face_q: np.ndarray
    dz_cum: np.ndarray
    banked: float
    mass_rel_error: float
    steps: int
    activations: int
    t: float
    courant: float

    def median_depth(self, where: slice) -> float:
        return float(np.median(self.depth[where]))

    def median_q(self, where: slice) -> float:
        return float(np.median(self.face_q[1:-1][where]))


def drive(rbs, fx: Flume, *, morphology: bool = True, end_s: float | None = None) -> FlumeRun:
    """The reference's fixture driver (bedwave.py:615-705) through librbs: the host loop
    asks compute_dt for each step, fills the span to the next landing with the
    scheduler's fill_span, injects the head inflow at the step midpoint, and at each
    interval past the warm-up advances the morphology process."""
    st = rbs.State.from_bed(fx.bed(), fx.dx, depth=fx.depth, manning=fx.manning)
    st.set_open_boundaries({"east": "open"})
    inj = rbs.inflow.InflowInjector(
        [rbs.inflow.Inflow((0, 0), [(0.0, fx.q * fx.dx), (1.0e9, fx.q * fx.dx)])], st)
    led = rbs.mb.MassLedger.from_state(st)
    morph = None
    end = (fx.warmup_s + fx.activations * fx.interval_s) if end_s is None else float(end_s)
    t, steps, acts = 0.0, 0, 0
    while t < end - EPS_T:
        dt = rbs.li.compute_dt(st, 0.7, DT_MAX)
        if not morphology:
            edge = end
        elif t < fx.warmup_s - EPS_T:
            edge = fx.warmup_s
        else:
            edge = fx.warmup_s + (acts + 1) * fx.interval_s
        dt = fill_span(min(edge, end) - t, dt)
        assert dt > 0.0
        if morphology and morph is None and t >= fx.warmup_s - EPS_T:
            rbs.sed.arm_sediment(st, fx.d50, fx.porosity)
            lo, hi = fx.bounds()
            morph = rbs.morph.MorphologyProcess(st, fx.interval_s, dz_lo=lo, dz_hi=hi)
        led.add_inflow(inj.apply(st, t, dt))
        rbs.li.step(st, dt, t=t)
        steps += 1
        t += dt
        if morph is not None and t >= fx.warmup_s + (acts + 1) * fx.interval_s - EPS_T:
            morph.advance(t, fx.interval_s)
            acts += 1
    led.record(st, t)
    sed = st.sediment
    run = FlumeRun(
        bed=st.z.numpy()[0].astype(np.float64), depth=st.h.numpy()[0].astype(np.float64),
        face_q=st.qx.numpy()[0].astype(np.float64),
        dz_cum=sed.dz_cum.numpy()[0] if sed is not None else np.zeros(fx.nx),
        banked=float(sed.dz_unapplied.numpy().sum()) if sed is not None else 0.0,
        mass_rel_error=led.max_rel_error, steps=steps, activations=acts, t=t,
        courant=morph.peak_courant if morph is not None else 0.0)
    st.close()
    return run
