"""Global mass-balance ledger (reference solver/core/massbalance.py:27-158).

Same contract and arithmetic as the reference: ``residual = inflow - outflow -
(V - V0)`` with the causal peak-volume denominator and ``MASS_GATE = 1e-6``.
What moves to the device is the float64 work the reference did on the host at
output cadence: ``V = dx^2 * sum f64(h)`` and the sink total are reduced on the GPU
in fixed order (deterministic), and a device-driven run also banks its rain and
point-source inflow there with Kahan compensation.  Inflow the caller adds on the
host (``add_inflow`` / ``add_rain_step``) keeps the reference's host Kahan sum.
"""

from __future__ import annotations

from dataclasses import dataclass, field

from rbs.core.state import State

MASS_GATE = 1.0e-6  # massbalance.py:36


@dataclass
class _Kahan:
    """Compensated float64 accumulator (massbalance.py:39-50)."""

    total: float = 0.0
    comp: float = 0.0

    def add(self, x: float) -> None:
        y = x - self.comp
        t = self.total + y
        self.comp = (t - self.total) - y
        self.total = t


@dataclass
class MassRecord:
    time: float
    volume: float
    inflow_cum: float
    outflow_cum: float
    residual: float
    rel_error: float


@dataclass
class MassLedger:
    cell_area: float
    v0: float
    _inflow: _Kahan = field(default_factory=_Kahan)
    _outflow: _Kahan = field(default_factory=_Kahan)
    series: list[MassRecord] = field(default_factory=list)
    _peak_v: float = 0.0
    _device_in0: float = 0.0
    _device_out0: float = 0.0

    @classmethod
    def from_state(cls, state: State) -> MassLedger:
        led = state.ledger()
        ledger = cls(cell_area=state.grid.cell_area, v0=led.volume)
        # Device flow totals at construction are the zero of this ledger.
        ledger._device_in0 = led.rain_in + led.point_in
        ledger._device_out0 = led.infil_out + led.edge_out + led.release_out
        ledger.record(state, 0.0)
        return ledger

    def add_inflow(self, volume: float) -> None:
        self._inflow.add(volume)

    def add_outflow(self, volume: float) -> None:
        self._outflow.add(volume)

    def add_rain_step(self, rain_m_s: float, dt: float, n_cells: int) -> None:
        self.add_inflow(rain_m_s * dt * self.cell_area * n_cells)

    def record(self, state: State, time: float) -> MassRecord:
        led = state.ledger()
        v = led.volume
        inflow = self._inflow.total + (led.rain_in + led.point_in - self._device_in0)
        outflow = self._outflow.total + (led.infil_out + led.edge_out + led.release_out
                                         - self._device_out0)
        residual = inflow - outflow - (v - self.v0)
        self._peak_v = max(self._peak_v, v)
        denom = max(abs(inflow), abs(v), self._peak_v, 1e-12)
        rec = MassRecord(time, v, inflow, outflow, residual, abs(residual) / denom)
        self.series.append(rec)
        return rec

    @property
    def max_rel_error(self) -> float:
        return max((r.rel_error for r in self.series), default=0.0)

    def as_attrs(self) -> dict:
        return {
            "mass_gate": MASS_GATE,
            "mass_max_rel_error": self.max_rel_error,
            "mass_balance_series": [r.__dict__.copy() for r in self.series],
        }


# --- Sediment (massbalance.py:161-306) ------------------------------------------------
SEDIMENT_GATE = 1.0e-11  # massbalance.py:178


@dataclass
class SedimentRecord:
    time: float
    bed_volume: float
    banked_volume: float
    gross_volume: float
    residual: float
    rel_error: float


@dataclass
class SedimentLedger:
    """Solid-volume balance of a morphology run: the domain is closed to bedload, so
    ``residual = -(bed + banked)`` against the causal peak of the gross displaced
    volume (massbalance.py:193-306)."""

    cell_area: float
    solid_fraction: float
    series: list[SedimentRecord] = field(default_factory=list)
    _peak_gross: float = 0.0

    @classmethod
    def from_state(cls, state: State) -> SedimentLedger:
        sed = getattr(state, "sediment", None)
        if sed is None:
            raise ValueError("SedimentLedger needs a state armed with arm_sediment()")
        led = cls(cell_area=state.grid.cell_area, solid_fraction=1.0 - sed.porosity)
        led.record(state, 0.0)
        return led

    def record(self, state: State, time: float) -> SedimentRecord:
        tot = state.sediment.totals()  # float64 device reductions of dz_cum, dz_unapplied
        solid = self.cell_area * self.solid_fraction
        bed = tot["sum"] * solid
        banked = tot["unapplied"] * solid
        gross = tot["sum_abs"] * solid
        residual = -(bed + banked)
        self._peak_gross = max(self._peak_gross, gross)
        denom = max(abs(banked), self._peak_gross, 1e-12)
        rec = SedimentRecord(float(time), bed, banked, gross, residual, abs(residual) / denom)
        self.series.append(rec)
        return rec

    @property
    def max_rel_error(self) -> float:
        return max((r.rel_error for r in self.series), default=0.0)

    def as_attrs(self) -> dict:
        return {"sediment_gate": SEDIMENT_GATE,
                "sediment_max_rel_error": self.max_rel_error,
                "sediment_balance_series": [r.__dict__.copy() for r in self.series]}
