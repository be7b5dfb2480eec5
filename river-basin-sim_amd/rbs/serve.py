"""Serving loop: continuous batching of scenario requests on one GPU.

Requests (a scenario + its simulated duration) arrive over time.  The server
admits an arrived request when its device footprint fits the HBM budget and a
batch slot is free, keeps its state resident on the GPU, and advances every
admitted request together in batched launches (:class:`rbs.ensemble.Batch`).
Requests that arrive while others are running join between launches; finished
requests leave and release their memory.  Morphology requests stop at each
activation for their Exner update and rejoin.  Every request's arithmetic is the
one it would get alone (the batch is bitwise-neutral), so results do not depend
on what else was being served.

Reported: time-to-result (arrival -> result on the host) p50/p99, scenarios/hour.
"""

from __future__ import annotations

import heapq
import time
from dataclasses import dataclass, field

import numpy as np

from rbs.driver import DeviceRun
from rbs.ensemble import Batch
from rbs.scheduler import MultiRateScheduler

HBM_BYTES = 288 * 10**9


@dataclass
class Request:
    """One scenario to simulate: inputs, duration and (optional) morphology."""

    rid: int
    arrival_s: float
    bed: np.ndarray
    manning: float | np.ndarray
    dx: float
    end_time: float
    rain: np.ndarray | None = None  # m/s field
    rain_duration: float = 0.0
    infil: np.ndarray | None = None
    d50: np.ndarray | float | None = None
    porosity: float = 0.4
    interval_s: float = 900.0
    inflows: list = field(default_factory=list)  # rbs.processes.inflow.Inflow
    open_edges: tuple[str, ...] = ("north", "south", "east", "west")
    kind: str = "basin"

    def footprint_bytes(self) -> int:
        """Device bytes the request's state will hold, as librbs computes them for the
        grid and the fields the request arms (rbs_state_footprint), before it exists."""
        import ctypes

        from rbs import _lib

        ny, nx = self.bed.shape
        arm = 0
        if self.rain is not None:
            arm |= _lib.ARM_COMP | _lib.ARM_RAIN
        if self.infil is not None:
            arm |= _lib.ARM_INFIL
        if self.d50 is not None:
            arm |= _lib.ARM_SED
        fields, total = ctypes.c_uint64(), ctypes.c_uint64()
        _lib.call("rbs_state_footprint", int(ny), int(nx), arm, ctypes.byref(fields),
                  ctypes.byref(total))
        return int(total.value)


@dataclass
class Result:
    rid: int
    kind: str
    cells: int
    arrival_s: float
    admitted_s: float
    done_s: float
    steps: int
    h: np.ndarray | None = None

    @property
    def time_to_result(self) -> float:
        return self.done_s - self.arrival_s


def build_state(req: Request, tune_kernel: bool = True):
    """The device state + run of one request.  ``tune_kernel`` picks the fastest LI
    kernel for the request's shape (a request run alone); a batch sets its own."""
    from rbs.core.sediment import arm_sediment
    from rbs.core.state import State
    from rbs.processes.inflow import InflowInjector
    from rbs.processes.morphology import MorphologyProcess

    st = State.from_bed(req.bed, req.dx, depth=0.0, manning=req.manning)
    rain_sum = 0.0
    if req.rain is not None:
        st.set_rain_field(req.rain)
        st.arm_source_compensation()
        rain_sum = float(req.rain.astype(np.float64).sum())
    if req.infil is not None:
        st.set_infiltration(req.infil)
    st.set_open_boundaries({e: "open" for e in req.open_edges})
    if req.inflows:
        InflowInjector(req.inflows, st)
    morph = None
    procs = []
    if req.d50 is not None:
        arm_sediment(st, req.d50, req.porosity)
        # a served request records its bed-change volumes; the Courant read-back (five
        # fields per activation) is left to runs that ask for it
        morph = MorphologyProcess(st, req.interval_s, diagnostics=False)
        procs = [morph.as_slow_process()]
    if tune_kernel:
        from rbs.autotune import tune

        tune(st)
    events = [req.rain_duration] if req.rain is not None else []
    for inf in req.inflows:
        events.extend(inf.breakpoints)
    sched = MultiRateScheduler(end_time=req.end_time, output_every=req.end_time, events=events,
                               processes=procs)
    run = DeviceRun(st, "local_inertial", sched, alpha=0.7, dt_max=30.0,
                    rain_is_field=req.rain is not None, rain_duration=req.rain_duration,
                    rain_field_sum=rain_sum)
    return st, run, morph


def plain(req: Request, keep_h: bool = True) -> Result:
    """Run one request alone the way the reference's run loop does (run.py:369-416):
    every tick asks ``compute_dt`` (a reduction and a device-to-host read), the host
    picks the step with ``fill_span``, applies point sources, launches one step, and
    runs slow processes at their activations.  The wall time covers building the state
    and the run, not the kernel-variant search."""
    from rbs.autotune import tune
    from rbs.core import local_inertial as li
    from rbs.processes.inflow import InflowInjector
    from rbs.scheduler import EPS_T

    t0 = time.perf_counter()
    st, run, morph = build_state(req, tune_kernel=False)
    inj = InflowInjector(req.inflows, st) if req.inflows else None
    t_build = time.perf_counter() - t0
    tune(st)  # the shape's fastest kernel, chosen outside the timed work
    t0 = time.perf_counter() - t_build
    steps = 0
    for tick in run.sched.ticks(lambda: li.compute_dt(st, 0.7, 30.0)):
        if inj is not None:
            inj.apply(st, tick.t0, tick.dt)
        raining = req.rain is not None and tick.t0 < req.rain_duration - EPS_T
        li.step(st, tick.dt, rain=0.0, rain_scale=1.0 if raining else 0.0, t=tick.t0)
        steps += 1
        for proc, elapsed in tick.due:
            proc.advance(tick.t1, elapsed)
    h = st.h.numpy() if keep_h else None
    t1 = time.perf_counter()
    st.close()
    return Result(req.rid, req.kind, req.bed.size, 0.0, 0.0, t1 - t0, steps, h)


def solo(req: Request, keep_h: bool = True) -> Result:
    """Run one request alone (device-driven, no batch) -- the reference result."""
    t0 = time.perf_counter()
    st, run, morph = build_state(req)
    steps = 0
    for seg in run.sched.segments():
        steps += run.run_until(seg.t_stop)
        for proc, elapsed in seg.due:
            proc.advance(seg.t_stop, elapsed)
    h = st.h.numpy() if keep_h else None
    t1 = time.perf_counter()
    st.close()
    return Result(req.rid, req.kind, req.bed.size, 0.0, 0.0, t1 - t0, steps, h)


@dataclass
class _Live:
    req: Request
    st: object
    run: DeviceRun
    morph: object
    segments: list
    seg_idx: int
    slot: int
    admitted_s: float
    bytes: int


class Server:
    """Continuous-batching server over one GPU."""

    def __init__(self, *, capacity: int = 256, hbm_budget: int = int(0.85 * HBM_BYTES),
                 chunk: int = 32, keep_results: bool = False, prep_threads: int = 2,
                 variants: tuple[int, int] | None = None):
        self.batch = Batch(capacity)
        if variants is not None:
            self.batch.set_variants(*variants)
        self.capacity = capacity
        self.budget = hbm_budget
        self.chunk = chunk
        self.keep = keep_results
        self.prep_threads = prep_threads
        self.resident = 0
        self.live: dict[int, _Live] = {}
        self.results: list[Result] = []
        self.rejected: list[int] = []
        self.peak_live = 0
        self.peak_bytes = 0
        # wall time by activity on the main thread: waiting for a first build with
        # nothing to step, batched stepping, stops (slow processes, results)
        self.t_admit = self.t_advance = self.t_stops = 0.0
        self.n_advance = 0
        self._n_prepared = 0

    def _reserve(self, req: Request):
        """Admission control: ``True`` when the request fits the HBM budget and a slot
        now (its bytes are reserved), ``False`` for a request that can never fit
        (refused), ``None`` when it must wait."""
        need = req.footprint_bytes()
        if need > self.budget:
            self.rejected.append(req.rid)  # can never fit: refuse, do not wait forever
            return False
        if self.resident + need > self.budget or len(self.live) + self._n_prepared >= self.capacity:
            return None
        self.resident += need
        self.peak_bytes = max(self.peak_bytes, self.resident)
        self._n_prepared += 1
        return True

    @staticmethod
    def _build(req: Request, now: float) -> "_Live":
        """Build a request's device state (runs on the preparation thread)."""
        st, run, morph = build_state(req, tune_kernel=False)
        segs = list(run.sched.segments())
        return _Live(req, st, run, morph, segs, 0, -1, now, req.footprint_bytes())

    def _join(self, lv: "_Live") -> None:
        lv.slot = self.batch.add(lv.run, lv.segments[0].t_stop)
        self.live[lv.slot] = lv
        self._n_prepared -= 1
        self.peak_live = max(self.peak_live, len(self.live))

    def _detach_stopped(self) -> list:
        """Members that reached a stop leave the batch (the tables follow at the next
        launch); their stop work runs while the batch steps the others."""
        out = []
        for slot in [s for s in list(self.live) if not self.batch.active(s)]:
            lv = self.live.pop(slot)
            self.batch.remove(slot)
            out.append(lv)
        return out

    def _handle_stops(self, stopped: list, now_fn) -> list:
        """Slow processes at the stop (on the member's own stream), or the result;
        returns the members that continue."""
        cont = []
        for lv in stopped:
            seg = lv.segments[lv.seg_idx]
            for proc, elapsed in seg.due:
                proc.advance(seg.t_stop, elapsed)
            lv.seg_idx += 1
            if lv.seg_idx < len(lv.segments):
                cont.append(lv)
                continue
            led = lv.st.ledger()
            h = lv.st.h.numpy() if self.keep else None
            done = now_fn()
            self.results.append(Result(lv.req.rid, lv.req.kind, lv.req.bed.size, lv.req.arrival_s,
                                       lv.admitted_s, done, int(led.steps), h))
            lv.st.close()
            self.resident -= lv.bytes
        return cont

    def _rejoin(self, lv: "_Live") -> None:
        lv.slot = self.batch.add(lv.run, lv.segments[lv.seg_idx].t_stop)
        self.live[lv.slot] = lv

    def serve(self, requests: list[Request], *, log=None) -> dict:
        """Serve requests whose ``arrival_s`` is wall-clock seconds after start."""
        pending = [(r.arrival_s, r.rid, r) for r in requests]
        heapq.heapify(pending)
        waiting: list[Request] = []
        t_start = time.perf_counter()
        now = lambda: time.perf_counter() - t_start  # noqa: E731
        last_log = 0.0
        # States are built on a preparation thread while the main thread keeps the
        # batch stepping; librbs handles are independent, and batch graph capture is
        # thread-local, so the two never wait on each other.
        from concurrent.futures import ThreadPoolExecutor

        pool = ThreadPoolExecutor(max_workers=self.prep_threads)
        building: list = []  # futures of _Live, in admission order
        try:
            stopped: list = []  # members detached at a stop, their stop work pending
            while pending or waiting or self.live or building or stopped:
                t = now()
                while pending and pending[0][0] <= t:
                    waiting.append(heapq.heappop(pending)[2])
                while waiting:  # FIFO admission against the HBM budget and the slots
                    ok = self._reserve(waiting[0])
                    if ok is None:
                        break
                    req = waiting.pop(0)
                    if ok:
                        building.append(pool.submit(self._build, req, now()))
                ready = [f for f in building if f.done()]
                building = [f for f in building if not f.done()]
                for f in ready:  # the batch is idle here: built members join
                    self._join(f.result())
                launched = False
                if self.live:
                    tb = time.perf_counter()
                    self.batch.launch(self.chunk)
                    self.n_advance += 1
                    launched = True
                # while the device steps: the stop work of members detached last round
                tc = time.perf_counter()
                cont = self._handle_stops(stopped, now) if stopped else []
                self.t_stops += time.perf_counter() - tc
                stopped = []
                if launched:
                    self.batch.collect()
                    self.t_advance += time.perf_counter() - tb
                    stopped = self._detach_stopped()
                elif building and not cont:
                    ta = time.perf_counter()
                    building[0].result()  # nothing to step: wait for the first build
                    self.t_admit += time.perf_counter() - ta
                elif pending and not waiting and not cont:
                    time.sleep(max(0.0, min(0.05, pending[0][0] - now())))
                for lv in cont:  # the batch is idle here: continuing members rejoin
                    self._rejoin(lv)
                if log and now() - last_log > 10.0:
                    last_log = now()
                    log(f"[serve] t={last_log:.0f}s done={len(self.results)} "
                        f"live={len(self.live)} building={len(building)} waiting={len(waiting)} "
                        f"pending={len(pending)} resident={self.resident / 1e9:.1f} GB")
        finally:
            pool.shutdown(wait=True)
        wall = now()
        ttr = np.array([r.time_to_result for r in self.results]) if self.results else np.zeros(1)
        bstats = self.batch.stats()
        self.batch.close()
        return {
            "completed": len(self.results),
            "rejected": len(self.rejected),
            "wall_s": wall,
            "scenarios_per_h": 3600.0 * len(self.results) / wall if wall > 0 else 0.0,
            "p50_s": float(np.percentile(ttr, 50)),
            "p99_s": float(np.percentile(ttr, 99)),
            "peak_live": self.peak_live,
            "peak_resident_gb": self.peak_bytes / 1e9,
            "time_idle_for_builds_s": self.t_admit,
            "time_stepping_s": self.t_advance,
            "time_stops_s": self.t_stops,
            "advances": self.n_advance,
            **{f"batch_{k}": v for k, v in bstats.items()},
        }


def poisson_arrivals(count: int, rate_per_s: float, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, count))


class InputPool:
    """Seeded base fields that requests crop windows from (generation stays cheap)."""

    def __init__(self, n: int = 4096, seed: int = 20261206):
        from rbs import terrain

        self.n = n
        self.dem = terrain.fbm_dem(n, n, hurst=0.8, relief=200.0, seed=seed)
        self.g = terrain.gaussian_field(n, n, 200.0, seed=seed + 1).astype(np.float32)
        self.relief2 = terrain.fbm_dem(n, n, hurst=0.8, relief=2.0, seed=seed + 2)

    def crop(self, a: np.ndarray, ny: int, nx: int, rng) -> np.ndarray:
        if ny > self.n or nx > self.n:  # periodic fields: tile to cover the window
            a = np.tile(a, (-(-ny // self.n), -(-nx // self.n)))
            i = int(rng.integers(0, a.shape[0] - ny + 1))
            j = int(rng.integers(0, a.shape[1] - nx + 1))
            return np.ascontiguousarray(a[i:i + ny, j:j + nx])
        i = int(rng.integers(0, self.n - ny + 1))
        j = int(rng.integers(0, self.n - nx + 1))
        return np.ascontiguousarray(a[i:i + ny, j:j + nx])


def basin_request(rid: int, arrival: float, pool: InputPool, rng, *, n_min=512, n_max=4096,
                  end_time=3600.0) -> Request:
    """Local-inertial basin: side log-uniform in [n_min, n_max], dx 10 m, lognormal
    rain (median U[5, 40] mm/h, sigma 0.6) for the first half, infiltration U[0, 10]
    mm/h, open edges."""
    from rbs.terrain import MM_H

    n = int(round(np.exp(rng.uniform(np.log(n_min), np.log(n_max)))))
    n = min(n, pool.n)
    bed = pool.crop(pool.dem, n, n, rng)
    median = rng.uniform(5.0, 40.0)
    rain = (median * np.exp(0.6 * pool.crop(pool.g, n, n, rng))).astype(np.float32) * np.float32(MM_H)
    infil = rng.uniform(0.0, 10.0, (n, n)).astype(np.float32) * np.float32(MM_H)
    return Request(rid=rid, arrival_s=float(arrival), bed=bed, manning=0.035, dx=10.0,
                   end_time=end_time, rain=rain, rain_duration=0.5 * end_time, infil=infil,
                   kind="basin")


def reach_request(rid: int, arrival: float, pool: InputPool, rng, *, end_time=3600.0,
                  scale: float = 1.0, factor: float | None = None) -> Request:
    """Alluvial reach, dx 2 m, length x width log-uniform between 1024x256 and
    4096x1024: a trapezoidal channel (8 m wide, 1.5 m deep) on slope 1e-3 with 2 m of
    fBm floodplain relief, an inflow hydrograph on the channel's upstream cells,
    Exner + MPM with d50 U[0.5, 4] mm, porosity 0.4, morphological factor 1, open
    downstream edge.  ``scale`` shrinks both sides (small test cases); ``factor``
    fixes the size factor in [1, 4] instead of drawing it."""
    from rbs.processes.inflow import Inflow

    f = (np.exp(rng.uniform(0.0, np.log(4.0))) if factor is None else float(factor)) * scale
    ny, nx = max(int(round(1024 * f)), 16), max(int(round(256 * f)), 16)
    dx = 2.0
    yy = np.arange(ny)[:, None] * np.ones((1, nx))
    xx = np.ones((ny, 1)) * np.arange(nx)[None, :]
    c = nx / 2.0
    dist = np.abs(xx - c) * dx
    chan = np.clip(1.5 - np.maximum(dist - 4.0, 0.0) * 0.5, 0.0, 1.5)  # 8 m bed, 1:2 banks
    bed = (1e-3 * dx * (ny - yy) - chan + pool.crop(pool.relief2, ny, nx, rng)
           * (dist > 12.0)).astype(np.float32)
    cells = [int(j) for j in np.arange(nx) if abs(j - c) * dx <= 4.0]
    q_peak = rng.uniform(5.0, 20.0)
    hg = [(0.0, 0.2 * q_peak), (0.4 * end_time, q_peak), (end_time, 0.2 * q_peak)]
    inflows = [Inflow((1, j), [(t, q / len(cells)) for t, q in hg]) for j in cells]
    d50 = float(rng.uniform(0.5e-3, 4e-3))
    return Request(rid=rid, arrival_s=float(arrival), bed=bed, manning=0.03, dx=dx,
                   end_time=end_time, d50=d50, porosity=0.4, interval_s=900.0,
                   inflows=inflows, open_edges=("south",), kind="reach")


def make_requests(count: int, *, rate_per_s: float = 4.0, seed: int = 20261206,
                  pool: InputPool | None = None, end_time: float = 3600.0,
                  basin_min: int = 512, basin_max: int = 4096,
                  reach_scale: float = 1.0) -> list[Request]:
    """The served mix: 70 % basins, 30 % alluvial reaches, Poisson arrivals."""
    if not 0 < basin_min <= basin_max:
        raise ValueError(f"basin sides need 0 < min <= max, got {basin_min}, {basin_max}")
    pool = pool or InputPool(seed=seed)
    rng = np.random.default_rng(seed)
    arrivals = poisson_arrivals(count, rate_per_s, seed + 7)
    out = []
    for k in range(count):
        if rng.uniform() < 0.7:
            out.append(basin_request(k, arrivals[k], pool, rng, n_min=basin_min,
                                     n_max=basin_max, end_time=end_time))
        else:
            out.append(reach_request(k, arrivals[k], pool, rng, end_time=end_time,
                                     scale=reach_scale))
    return out
