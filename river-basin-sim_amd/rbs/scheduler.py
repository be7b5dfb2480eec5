"""Multi-rate scheduler: the single float64 simulated clock (reference
solver/scheduler.py:1-254).

The contract is the reference's: sync points (outputs, forcing events, slow-process
activations, end) are never crossed; a span is filled with ``ceil(span/dt_raw)``
equal steps, never longer than ``dt_raw``; slow processes are handed the exact
elapsed simulated time.  ``ticks`` drives a host loop one step at a time (the
reference's seam); ``segments`` hands a device-driven run the sync points it must
stop at, between which librbs chooses the steps itself with the same arithmetic
(device ``fill_span``).
"""

from __future__ import annotations

import math
from collections.abc import Callable, Iterable, Iterator, Sequence
from dataclasses import dataclass, field

EPS_T = 1e-6  # seconds: two clock readings this close are the same instant (scheduler.py:45)


def fill_span(span: float, dt_raw: float) -> float:
    """The step that fills ``span`` (the time left to the next sync point) with equal
    steps no longer than the CFL step ``dt_raw``, in float64 (scheduler.py:48-72).

    A span shorter than one step is taken whole; otherwise ``k = ceil(span/dt_raw)``
    equal steps, one more when the division's rounding leaves ``span/k`` a hair above
    ``dt_raw``.  An infinite span (nothing ahead) leaves ``dt_raw`` unchanged."""
    if not math.isfinite(span):
        return dt_raw
    if span <= dt_raw:
        return span
    k = math.ceil(span / dt_raw)
    dt = span / k
    return span / (k + 1) if dt > dt_raw else dt


@dataclass(frozen=True)
class SlowProcess:
    """Something the run advances at its own cadence, ``interval`` seconds of simulated
    time apart, always on a sync point (scheduler.py:75-99).  ``advance(t, elapsed)``
    receives the clock and the exact simulated time since its previous activation."""

    name: str
    interval: float
    advance: Callable[[float, float], object]

    def __post_init__(self) -> None:
        if self.interval <= 0:
            raise ValueError(f"slow process '{self.name}': interval must be > 0")

    def activations(self, end_time: float) -> list[float]:
        """Whole multiples of the interval up to ``end_time`` (inclusive within EPS_T)."""
        count = int((end_time + EPS_T) // self.interval)
        return [self.interval * k for k in range(1, count + 1)]


@dataclass(frozen=True)
class Tick:
    """A fast step t0 -> t1 = t0 + dt, whether an output lands on t1, and the slow
    processes due there with their elapsed times (scheduler.py:102-116)."""

    index: int
    t0: float
    dt: float
    t1: float
    is_output: bool
    due: tuple[tuple[SlowProcess, float], ...] = ()


@dataclass(frozen=True)
class Segment:
    """A stretch a device-driven run covers without the host: it ends on ``t_stop``,
    where the host writes an output (``is_output``) and advances the ``due`` processes
    -- what :meth:`MultiRateScheduler.ticks` reports for the tick landing there."""

    t_stop: float
    is_output: bool
    due: tuple[tuple[SlowProcess, float], ...] = ()


class _Landings:
    """What happens when the clock lands on a time: the output cadence and each slow
    process's pending activations, consumed in order as the clock passes them."""

    def __init__(self, sched: "MultiRateScheduler"):
        self.every = sched.output_every
        self.end = sched.end_time
        self.next_output = sched.output_every
        self.pending = [(p, list(p.activations(sched.end_time))) for p in sched.processes]
        self.last = [0.0] * len(self.pending)

    def at(self, t1: float) -> tuple[bool, tuple[tuple[SlowProcess, float], ...]]:
        is_output = t1 >= self.next_output - EPS_T and self.next_output <= self.end + EPS_T
        if is_output:
            self.next_output += self.every
        due = []
        for k, (proc, times) in enumerate(self.pending):
            passed = 0
            while passed < len(times) and t1 >= times[passed] - EPS_T:
                passed += 1
            if passed:
                del times[:passed]
                due.append((proc, t1 - self.last[k]))
                self.last[k] = t1
        return is_output, tuple(due)


@dataclass
class MultiRateScheduler:
    """The float64 clock of one run (scheduler.py:119-246).

    Its sync points are every output time, every forcing event, every slow-process
    activation and the end time; no step crosses one.  ``ticks`` paces a host loop one
    step at a time, the reference's seam; ``segments`` gives a device-driven run the
    sync points where the host has work, and the device fills the spans between them
    with the same ``fill_span`` arithmetic."""

    end_time: float
    output_every: float
    events: Sequence[float] = ()
    processes: Sequence[SlowProcess] = ()
    _sync: list[float] = field(init=False, repr=False)

    def __post_init__(self) -> None:
        if self.end_time <= 0:
            raise ValueError(f"end_time must be > 0, got {self.end_time}")
        if self.output_every <= 0:
            raise ValueError(f"output_every must be > 0, got {self.output_every}")
        sync = [*self.output_times, *self.events]
        for proc in self.processes:
            sync += proc.activations(self.end_time)
        sync.append(self.end_time)
        self._sync = sync

    @property
    def n_frames(self) -> int:
        """Frames a run writes, the one at t = 0 included."""
        return int(round(self.end_time / self.output_every)) + 1

    @property
    def output_times(self) -> list[float]:
        return [self.output_every * k for k in range(1, self.n_frames)]

    @property
    def sync_times(self) -> list[float]:
        return list(self._sync)

    def sorted_sync_times(self) -> list[float]:
        """The sync table ascending, as the device scans it for the next stop."""
        return sorted(self._sync)

    def next_sync(self, t: float) -> float:
        """The first sync point after ``t`` (beyond EPS_T), or infinity."""
        ahead = [s for s in self._sync if s > t + EPS_T]
        return min(ahead) if ahead else float("inf")

    def ticks(self, dt_fn: Callable[[], float]) -> Iterator[Tick]:
        """Fast steps to ``end_time``: each asks ``dt_fn`` for the CFL step and fills the
        span to the next sync point with it (scheduler.py:171-246)."""
        land = _Landings(self)
        t, index = 0.0, 0
        while t < self.end_time - EPS_T:
            dt_raw = float(dt_fn())
            if dt_raw <= 0.0:
                raise RuntimeError(f"non-positive timestep {dt_raw!r} at t={t}")
            dt = fill_span(self.next_sync(t) - t, dt_raw)
            if dt <= 0.0:
                raise RuntimeError(f"non-positive timestep {dt!r} at t={t}")
            t1 = t + dt
            is_output, due = land.at(t1)
            yield Tick(index=index, t0=t, dt=dt, t1=t1, is_output=is_output, due=due)
            t, index = t1, index + 1

    def segments(self, stop_at_events: bool = False) -> Iterator[Segment]:
        """The sync points where the host has work, in order: outputs, slow-process
        activations, the end (and forcing events too with ``stop_at_events``; without
        it the device handles them itself -- rain switching off, hydrograph and stage
        knots -- as sync points of its step choice)."""
        stops = {*self.output_times, self.end_time}
        for proc in self.processes:
            stops.update(proc.activations(self.end_time))
        if stop_at_events:
            stops.update(e for e in self.events if 0.0 < e <= self.end_time)
        land = _Landings(self)
        t = 0.0
        for t1 in sorted(stops):
            if t1 <= t + EPS_T:
                continue
            is_output, due = land.at(t1)
            yield Segment(t_stop=t1, is_output=is_output, due=due)
            t = t1
            if t >= self.end_time - EPS_T:
                break


def merge_events(*groups: Iterable[float]) -> list[float]:
    """One event list from several groups of forcing breakpoints (scheduler.py:249-254)."""
    return [float(x) for g in groups for x in g]
