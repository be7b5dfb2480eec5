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

EPS_T = 1e-6  # scheduler.py:45


def fill_span(span: float, dt_raw: float) -> float:
    """The step to take when ``span`` remains to the next sync point (scheduler.py:48-72)."""
    if not math.isfinite(span):
        return dt_raw
    if span <= dt_raw:
        return span
    n = math.ceil(span / dt_raw)
    dt = span / n
    if dt > dt_raw:
        dt = span / (n + 1)
    return dt


@dataclass(frozen=True)
class SlowProcess:
    """A process advanced at sync points on its own cadence (scheduler.py:75-99)."""

    name: str
    interval: float
    advance: Callable[[float, float], object]

    def __post_init__(self) -> None:
        if self.interval <= 0:
            raise ValueError(f"slow process '{self.name}': interval must be > 0")

    def activations(self, end_time: float) -> list[float]:
        n = int((end_time + EPS_T) // self.interval)
        return [self.interval * k for k in range(1, n + 1)]


@dataclass(frozen=True)
class Tick:
    """One fast step from ``t0`` to ``t1`` (scheduler.py:102-116)."""

    index: int
    t0: float
    dt: float
    t1: float
    is_output: bool
    due: tuple[tuple[SlowProcess, float], ...] = ()


@dataclass(frozen=True)
class Segment:
    """A device-driven stretch ending on a sync point where the host has work.

    ``t_stop`` is the sync time the device clock must land on; ``is_output`` and
    ``due`` are what the host does there (the same bookkeeping ``ticks`` does on the
    tick that lands on ``t_stop``).
    """

    t_stop: float
    is_output: bool
    due: tuple[tuple[SlowProcess, float], ...] = ()


@dataclass
class MultiRateScheduler:
    """The clock of one run (scheduler.py:119-246)."""

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
        self._sync = list(self.output_times) + list(self.events)
        for p in self.processes:
            self._sync.extend(p.activations(self.end_time))
        self._sync.append(self.end_time)

    @property
    def n_frames(self) -> int:
        return int(round(self.end_time / self.output_every)) + 1

    @property
    def output_times(self) -> list[float]:
        return [self.output_every * k for k in range(1, self.n_frames)]

    @property
    def sync_times(self) -> list[float]:
        return list(self._sync)

    def sorted_sync_times(self) -> list[float]:
        """Sync times ascending (the device's ``next_sync`` scans this table)."""
        return sorted(self._sync)

    def next_sync(self, t: float) -> float:
        future = [e for e in self._sync if e > t + EPS_T]
        return min(future) if future else float("inf")

    def ticks(self, dt_fn: Callable[[], float]) -> Iterator[Tick]:
        """One :class:`Tick` per fast step until ``end_time`` (scheduler.py:171-246)."""
        next_output = self.output_every
        next_act = {p.name: (p, list(p.activations(self.end_time))) for p in self.processes}
        last_act = dict.fromkeys(next_act, 0.0)
        t = 0.0
        index = 0
        while t < self.end_time - EPS_T:
            dt_raw = float(dt_fn())
            if dt_raw <= 0.0:
                raise RuntimeError(f"non-positive timestep {dt_raw!r} at t={t}")
            dt = fill_span(self.next_sync(t) - t, dt_raw)
            if dt <= 0.0:
                raise RuntimeError(f"non-positive timestep {dt!r} at t={t}")
            t1 = t + dt
            is_output = t1 >= next_output - EPS_T and next_output <= self.end_time + EPS_T
            if is_output:
                next_output += self.output_every
            due: list[tuple[SlowProcess, float]] = []
            for name, (proc, times) in next_act.items():
                fired = False
                while times and t1 >= times[0] - EPS_T:
                    times.pop(0)
                    fired = True
                if fired:
                    due.append((proc, t1 - last_act[name]))
                    last_act[name] = t1
            yield Tick(index=index, t0=t, dt=dt, t1=t1, is_output=is_output, due=tuple(due))
            t = t1
            index += 1

    def segments(self, stop_at_events: bool = False) -> Iterator[Segment]:
        """Sync points where the host has work, in order, with their bookkeeping.

        The host stops the device at every output time and slow-process activation
        (and at every event if ``stop_at_events``); forcing events the device
        evaluates itself (rain on/off, hydrograph knots, stage knots) stay sync points
        of the device's step choice without a host stop.  The ``is_output``/``due``
        arithmetic is the one :meth:`ticks` applies to the tick landing there.
        """
        stops = set(self.output_times) | {self.end_time}
        for p in self.processes:
            stops.update(p.activations(self.end_time))
        if stop_at_events:
            stops.update(e for e in self.events if 0.0 < e <= self.end_time)
        next_output = self.output_every
        next_act = {p.name: (p, list(p.activations(self.end_time))) for p in self.processes}
        last_act = dict.fromkeys(next_act, 0.0)
        t = 0.0
        for t1 in sorted(stops):
            if t1 <= t + EPS_T:
                continue
            is_output = t1 >= next_output - EPS_T and next_output <= self.end_time + EPS_T
            if is_output:
                next_output += self.output_every
            due: list[tuple[SlowProcess, float]] = []
            for name, (proc, times) in next_act.items():
                fired = False
                while times and t1 >= times[0] - EPS_T:
                    times.pop(0)
                    fired = True
                if fired:
                    due.append((proc, t1 - last_act[name]))
                    last_act[name] = t1
            yield Segment(t_stop=t1, is_output=is_output, due=tuple(due))
            t = t1
            if t >= self.end_time - EPS_T:
                break


def merge_events(*groups: Iterable[float]) -> list[float]:
    """Flatten forcing-breakpoint groups into one event list (scheduler.py:249-254)."""
    out: list[float] = []
    for g in groups:
        out.extend(float(x) for x in g)
    return out
