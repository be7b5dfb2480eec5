"""Benchmark: local-inertial rain-on-terrain time step on one MI355X.

Workload (the headline configuration): 4096 x 4096 FP32 grid, dx = 10 m, seeded
fBm DEM (H 0.8, 200 m relief) with D8-carved channels 1-3 m deep, Voronoi Manning
n in [0.02, 0.10], lognormal rain field (median 15 mm/h, sigma 0.6, 2 km
correlation), infiltration U[0, 10] mm/h, all edges open, Kahan-compensated
sources.  The state is spun up device-driven for ``--spinup`` simulated seconds of
storm (not timed), then W warm-up steps, then K timed steps, all device-driven:
each step is one fused LI kernel + one finalize kernel chosen by the device CFL
reduction and replayed from a captured HIP graph.

Prints one JSON line.  ``value`` = cell-updates / s of the timed region (inputs
resident in HBM).  ``roofline.achieved`` = 48 algorithmic bytes per cell-update x
cells / the fused kernel's mean duration (HIP events on the library's stream).
``cpu_baseline`` times the C restatement oracle on the host cores on a bounded
sample of the same state.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "river-basin-sim_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

import numpy as np  # noqa: E402

METRIC = "Gcell-updates/s + %HBM peak; ensemble scenarios/h & p99 latency vs plain"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector peak (datasheet)
BYTES_PER_CELL = 48  # h r+w, qx r+w, qy r+w, z, n, rain, infil, h_comp r+w


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def build_state(n: int, seed: int):
    from rbs import terrain
    from rbs.core.state import State

    t0 = time.time()
    d = terrain.rain_on_terrain(n, dx=10.0, seed=seed)
    log(f"[bench] generated {n}x{n} inputs in {time.time() - t0:.1f} s")
    st = State.from_bed(d["bed"], d["dx"], depth=0.0, manning=d["manning"])
    st.set_rain_field(d["rain"])
    st.set_infiltration(d["infil"])
    st.set_open_boundaries({e: "open" for e in ("north", "south", "east", "west")})
    st.arm_source_compensation()
    return st, d


def host_threads() -> int:
    """The host threads this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS when that is set (a shared host's share of its cores)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(cap))) if cap.isdigit() and int(cap) > 0 else max(1, n)


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(st, d, dt: float, threads: int, budget_s: float) -> dict:
    """Time the oracle (C restatement, -O3, OpenMP) on a copy of the current state."""
    from oracle import oracle as orc

    o = orc.State(d["bed"], d["dx"], depth=st.h.numpy(), manning=d["manning"])
    o.qx[...] = st.qx.numpy()
    o.qy[...] = st.qy.numpy()
    o.set_rain_field(d["rain"])
    o.set_infiltration(d["infil"])
    o.set_open_boundaries({e: "open" for e in ("north", "south", "east", "west")})
    o.arm_source_compensation()
    o.h_comp[...] = st.h_comp.numpy()
    cells = o.ny * o.nx
    out = {}
    for nthr in sorted({1, threads}):
        orc.set_threads(nthr)
        orc.set_pow("libm")  # the reference evaluates wp.pow with the C library's powf
        orc.li_run_fixed(o, dt, 1, rain_scale=1.0)  # warm caches / threads
        steps, el = 0, 0.0
        t0 = time.perf_counter()
        while el < budget_s / (2 if nthr == 1 else 4) or steps < 2:
            orc.li_run_fixed(o, dt, 1, rain_scale=1.0)
            steps += 1
            el = time.perf_counter() - t0
        out[nthr] = (cells * steps / el / 1e9, steps, el)
        log(f"[bench] cpu oracle {nthr} thread(s): {steps} steps in {el:.1f} s "
            f"-> {out[nthr][0]:.4f} Gcell/s")
    v, steps, el = out[threads]
    return {
        "value": round(v, 5),
        "unit": "Gcell-updates/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{steps} steps of the {o.ny}x{o.nx} spun-up state, {el:.1f} s; "
                  f"1 thread: {out[1][0]:.5f} Gcell-updates/s over {out[1][1]} steps",
        "value_1thread": round(out[1][0], 6),
        "cpu_model": cpu_model(),
        "cpus_visible": os.cpu_count(),
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=4096, help="grid side (cells)")
    ap.add_argument("--spinup", type=float, default=3600.0, help="simulated s before timing")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--variant", type=int, default=-1,
                    help="LI kernel variant (-1: the library's choice for the grid)")
    ap.add_argument("--tune", action="store_true", help="autotune the LI kernel variant")
    ap.add_argument("--hllc-variant", type=int, default=-1,
                    help="HLLC stage-kernel variant (-1: the library's default)")
    ap.add_argument("--hllc-hours", type=float, default=4.0,
                    help="hllc: simulated hours of the whole-flood run (0: the early window only)")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline only: skip the ensemble and serving figures")
    ap.add_argument("--serve-sample", type=int, default=120,
                    help="requests in the default line's serving sample")
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", default="li",
                    choices=["li", "ensemble", "serve", "hllc", "copy", "channels", "output",
                             "reservoir", "reach"])
    ap.add_argument("--host-loop", action="store_true",
                    help="reach: drive the runs from the host tick loop (no graphs)")
    ap.add_argument("--sweep", action="store_true",
                    help="also time every LI kernel variant on the spun-up state")
    ap.add_argument("--scenarios", type=int, default=256, help="ensemble size")
    ap.add_argument("--requests", type=int, default=2000, help="served requests")
    ap.add_argument("--rate", type=float, default=4.0, help="served arrivals per second")
    ap.add_argument("--chunk", type=int, default=64, help="serve: batched steps per advance")
    ap.add_argument("--sed-variant", type=int, default=-1,
                    help="serve: batch LI variant for bedload members (-1: library default)")
    ap.add_argument("--capacity-sample", type=int, default=120,
                    help="serve: requests replayed as a burst and as plain launches")
    ap.add_argument("--sim-hours", type=float, default=1.0, help="simulated time per scenario")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path "
              "(a scenario would shard by row strips, one per GPU)", file=sys.stderr)
        return 2

    from rbs import _lib
    from rbs.driver import DeviceRun
    from rbs.scheduler import MultiRateScheduler

    _lib.load()
    if args.workload == "ensemble":
        return bench_ensemble(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "hllc":
        return bench_hllc(args)
    if args.workload == "copy":
        return bench_copy(args)
    if args.workload == "channels":
        return bench_channels(args)
    if args.workload == "output":
        return bench_output(args)
    if args.workload == "reach":
        return bench_reach(args)
    if args.workload == "reservoir":
        return bench_reservoir(args)
    st, d = build_state(args.n, args.seed)
    if args.variant >= 0:
        _lib.call("rbs_set_li_variant", st._handle, args.variant)
        variant = args.variant
    elif args.tune:
        from rbs.autotune import tune

        variant = tune(st)
        log(f"[bench] autotuned LI variant {variant}")
    else:  # the kernel a state of this shape starts on: the same one every run
        v = ctypes.c_int32()
        _lib.call("rbs_get_li_variant", st._handle, ctypes.byref(v))
        variant = int(v.value)
        log(f"[bench] LI variant {variant} (the library's choice for this shape)")
    cells = args.n * args.n
    rain_sum = float(d["rain"].astype(np.float64).sum())
    end_time = 10800.0
    sched = MultiRateScheduler(end_time=end_time, output_every=600.0, events=[7200.0])
    run = DeviceRun(st, "local_inertial", sched, alpha=0.7, dt_max=30.0, rain_is_field=True,
                    rain_duration=7200.0, rain_field_sum=rain_sum)

    t0 = time.time()
    n_spin = 0
    for seg in sched.segments():
        if seg.t_stop > args.spinup + 1e-6:
            break
        n_spin += run.run_until(seg.t_stop)
    log(f"[bench] spin-up to t={st.ledger().t:.0f} s: {n_spin} steps in {time.time() - t0:.1f} s")
    run.run_until(end_time, max_steps=args.warmup)
    st.sync()

    led0 = st.ledger()
    t0 = time.perf_counter()
    taken = run.run_until(end_time, max_steps=args.steps)
    st.sync()
    el = time.perf_counter() - t0
    led1 = st.ledger()
    if taken != args.steps:
        log(f"[bench] note: {taken} steps taken (end of run reached)")
    value = cells * taken / el / 1e9
    ms_step = el * 1e3 / max(taken, 1)
    log(f"[bench] timed {taken} steps in {el:.3f} s: {value:.2f} Gcell/s, {ms_step:.4f} ms/step; "
        f"t {led0.t:.1f} -> {led1.t:.1f} s, h_max {led1.speed_max:.3f} m")

    # per-kernel duration of the fused LI kernel (HIP events on the library stream)
    dt_k = float(led1.dt) if led1.dt > 0 else 1.0
    avg = ctypes.c_double()
    _lib.call("rbs_time_li_kernel", st._handle, args.kernel_reps, dt_k, 0.0, 1.0,
              ctypes.byref(avg))
    kern_ms = float(avg.value)
    achieved = BYTES_PER_CELL * cells / (kern_ms * 1e-3) / 1e9
    log(f"[bench] fused LI kernel mean {kern_ms * 1e3:.1f} us -> {achieved:.0f} GB/s "
        f"({100 * achieved / HBM_PEAK_GBS:.1f}% of 8 TB/s)")
    sweep = None
    if args.sweep:
        sweep = variant_sweep(st, dt_k, args.kernel_reps, cells, args.variant)
    # HBM bytes per launch of the variant that ran, from the committed FETCH_SIZE /
    # WRITE_SIZE summary (null for a variant or grid it does not cover)
    traffic = None
    tf = ROOT / "profiles" / "li_rain_4096_traffic.json"
    if tf.exists() and args.n == 4096:
        entry = json.loads(tf.read_text()).get("variants", {}).get(str(variant))
        traffic = entry.get("bytes_per_launch") if entry else None

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(st, d, dt_k, host_threads(), args.cpu_budget)

    mass = None
    try:
        inflow = led1.rain_in
        outflow = led1.infil_out + led1.edge_out
        mass = abs(inflow - outflow - (led1.volume - 0.0)) / max(inflow + outflow, 1e-12)
    except Exception:
        pass

    extras = {}
    if not args.no_extras:
        st.close()  # the headline state's HBM back before the workloads below
        extras["ensemble"] = ensemble_figures(args)
        extras["serving"] = serve_figures(args, args.serve_sample, min(60, args.capacity_sample))
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": taken,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"li_rain_on_terrain_{args.n}",
            "grid": f"{args.n}x{args.n}",
            "dx_m": 10.0,
            "scheme": "local_inertial",
            "sources": "lognormal rain field (Kahan) + infiltration, open edges",
            "spinup_s": args.spinup,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel_us": round(kern_ms * 1e3, 2),
            "bytes_per_cell": BYTES_PER_CELL,
        },
        "cpu_baseline": cpu,
        "mass_closure_of_throughput": mass,
    }
    if sweep:
        out["variant_sweep"] = sweep
    out.update(extras)
    print(json.dumps(out), flush=True)
    return 0


def variant_sweep(st, dt: float, reps: int, cells: int, keep: int) -> list:
    """Mean fused-kernel duration of every LI variant on the same (advancing) state."""
    from rbs import _lib

    count = ctypes.c_int32()
    _lib.call("rbs_li_variant_count", ctypes.byref(count))
    rows = []
    for v in range(count.value):
        bx, by, fam, dflt = (ctypes.c_int32() for _ in range(4))
        _lib.call("rbs_li_variant_info", v, ctypes.byref(bx), ctypes.byref(by),
                  ctypes.byref(fam), ctypes.byref(dflt))
        _lib.call("rbs_set_li_variant", st._handle, v)
        avg = ctypes.c_double()
        _lib.call("rbs_time_li_kernel", st._handle, reps, dt, 0.0, 1.0, ctypes.byref(avg))
        us = avg.value * 1e3
        gbs = BYTES_PER_CELL * cells / (us * 1e-6) / 1e9
        fam_name = ("lds_tile", "strip", "strip2", "rows")[fam.value]
        rows.append({"variant": v, "family": fam_name, "tile": f"{bx.value}x{by.value}",
                     "default": bool(dflt.value), "kernel_us": round(us, 2),
                     "frac": round(gbs / HBM_PEAK_GBS, 4)})
        log(f"[bench] variant {v} {fam_name} {bx.value}x{by.value}: {us:.1f} us, "
            f"{100 * gbs / HBM_PEAK_GBS:.1f}% of 8 TB/s")
    if keep >= 0:
        _lib.call("rbs_set_li_variant", st._handle, keep)
    else:
        for r in rows:
            if r["default"]:
                _lib.call("rbs_set_li_variant", st._handle, r["variant"])
    return rows


def bench_copy(args) -> int:
    """A plain device copy of known size, for calibrating the HBM byte counters
    (FETCH_SIZE / WRITE_SIZE) that the traffic figure is scaled by."""
    import torch

    n = 256 * 1024 * 1024  # 1 GiB of float32
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    launches = 5
    for _ in range(launches):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"workload": "copy_calibration", "bytes_read_per_launch": 4 * n,
                      "bytes_written_per_launch": 4 * n, "launches": launches}), flush=True)
    return 0


def ensemble_figures(args) -> dict:
    """The batched ensemble against the same scenarios run one after another: as plain
    per-scenario launches (the reference's tick loop over compute_dt/step) and as
    device-driven graph-replayed runs.  Every scenario runs its shape's tuned kernel
    (tuned once per scenario, outside the timings); every result is compared bitwise."""
    from rbs import _lib
    from rbs.autotune import tune
    from rbs.core import local_inertial as li
    from rbs.core.state import State
    from rbs.driver import DeviceRun
    from rbs.ensemble import Batch, ensemble_inputs
    from rbs.scheduler import MultiRateScheduler

    t0 = time.time()
    members = ensemble_inputs(args.scenarios, seed=20261205)
    cells = sum(m.bed.size for m in members)
    log(f"[ensemble] {len(members)} scenarios, {cells / 1e6:.1f} M cells, generated in "
        f"{time.time() - t0:.0f} s")
    end = 3600.0 * args.sim_hours
    variants: dict[int, int] = {}

    def state(k):
        m = members[k]
        st = State.from_bed(m.bed, m.dx, depth=0.0, manning=m.manning)
        st.set_rain_field(m.rain)
        st.set_open_boundaries({e: "open" for e in ("north", "south", "east", "west")})
        st.arm_source_compensation()
        if k not in variants:
            variants[k] = tune(st)
        else:
            _lib.call("rbs_set_li_variant", st._handle, variants[k])
        return st

    def run(k):
        m = members[k]
        sched = MultiRateScheduler(end_time=end, output_every=end)
        return DeviceRun(state(k), "local_inertial", sched, alpha=0.7, dt_max=30.0,
                         rain_is_field=True, rain_duration=end,
                         rain_field_sum=float(m.rain.astype(np.float64).sum()))

    # plain per-scenario launches: the reference's tick loop (run.py:369-388) over the
    # scheme seam -- compute_dt (one reduction + a device-to-host read), fill_span on
    # the host, one step launch -- one scenario after another
    t_plain, plain_steps, h_plain = 0.0, 0, []
    for k in range(len(members)):
        st = state(k)
        sched = MultiRateScheduler(end_time=end, output_every=end)
        st.sync()
        t0 = time.perf_counter()
        for tick in sched.ticks(lambda: li.compute_dt(st, 0.7, 30.0)):
            li.step(st, tick.dt, rain=0.0, rain_scale=1.0 if tick.t0 < end - 1e-6 else 0.0,
                    t=tick.t0)
            plain_steps += 1
        st.sync()
        t_plain += time.perf_counter() - t0
        h_plain.append(st.h.numpy())
        st.close()
    log(f"[ensemble] plain per-scenario launches: {t_plain:.2f} s, {plain_steps} steps")
    # sequential device-driven runs: one scenario at a time, steps replayed from a graph
    t_seq, seq_steps, same = 0.0, 0, True
    for k in range(len(members)):
        r = run(k)
        r.state.sync()
        t0 = time.perf_counter()
        seq_steps += r.run_until(end)
        r.state.sync()
        t_seq += time.perf_counter() - t0
        same = same and np.array_equal(r.state.h.numpy(), h_plain[k])
        r.state.close()
    log(f"[ensemble] sequential: {t_seq:.2f} s GPU wall, {seq_steps} steps")
    runs = [run(k) for k in range(len(members))]
    for r in runs:
        r.state.sync()
    b = Batch(len(runs))
    t0 = time.perf_counter()
    slots = [b.add(r, end) for r in runs]
    b.run_to_completion(64)
    t_bat = time.perf_counter() - t0
    for sl in slots:
        b.remove(sl)
    b.close()
    same = same and all(np.array_equal(r.state.h.numpy(), h) for r, h in zip(runs, h_plain))
    bat_steps = sum(int(r.state.ledger().steps) for r in runs)
    updates = sum(m.bed.size * int(r.state.ledger().steps) for m, r in zip(members, runs))
    for r in runs:
        r.state.close()
    n = len(members)
    log(f"[ensemble] batched: {t_bat:.2f} s wall, {bat_steps} steps, identical={same}")
    return {
        "scenarios": n,
        "scenarios_per_h": round(3600.0 * n / t_bat, 1),
        "wall_s": round(t_bat, 3),
        "sides": "log-uniform 256..2048, dx 10 m", "simulated_s": end, "cells_total": cells,
        "plain_launches_scenarios_per_h": round(3600.0 * n / t_plain, 1),
        "speedup_vs_plain_launches": round(t_plain / t_bat, 3),
        "sequential_device_driven_scenarios_per_h": round(3600.0 * n / t_seq, 1),
        "speedup_vs_sequential_device_driven": round(t_seq / t_bat, 3),
        "bitwise_identical_batched_sequential_plain": bool(same),
        "batched_cell_updates_per_s": round(updates / t_bat / 1e9, 3),
        "steps": bat_steps,
    }


def bench_ensemble(args) -> int:
    """The 256-scenario ensemble as its own JSON line: scenarios/h of the batch (ensemble_figures)."""
    e = ensemble_figures(args)
    out = {
        "metric": METRIC,
        "value": e["scenarios_per_h"],
        "unit": "scenarios/h",
        "n_gpus": 1,
        "steps": e["steps"],
        "warmup": 0,
        "ms_per_step": round(1e3 * e["wall_s"] / max(1, e["steps"] / e["scenarios"]), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"li_ensemble_{e['scenarios']}", "sides": e["sides"],
                   "dx_m": 10.0, "simulated_s": e["simulated_s"], "cells_total": e["cells_total"]},
        **{k: v for k, v in e.items() if k not in ("scenarios_per_h", "steps")},
    }
    print(json.dumps(out), flush=True)
    return 0


def serve_figures(args, requests: int, capacity_sample: int) -> dict:
    """The continuous-batching server on ``requests`` Poisson arrivals of the served mix,
    with p50/p99 time-to-result, and its capacity on the first ``capacity_sample`` of
    them as one burst against plain per-request launches."""
    from rbs.serve import InputPool, Server, make_requests

    t0 = time.time()
    pool = InputPool(seed=20261206)
    reqs = make_requests(requests, rate_per_s=args.rate, seed=20261206, pool=pool,
                         end_time=3600.0 * args.sim_hours)
    log(f"[serve] {len(reqs)} requests generated in {time.time() - t0:.0f} s")
    variants = None
    if args.variant >= 0 or args.sed_variant >= 0:
        vd, vs = default_batch_variants()
        variants = (args.variant if args.variant >= 0 else vd,
                    args.sed_variant if args.sed_variant >= 0 else vs)
    srv = Server(capacity=256, chunk=args.chunk, variants=variants)
    stats = srv.serve(reqs, log=log)
    log(f"[serve] {stats}")
    cap = (serve_capacity(reqs[:capacity_sample], args.chunk, variants)
           if capacity_sample > 0 else None)
    kinds = {}
    for r in srv.results:
        kinds.setdefault(r.kind, []).append(r.time_to_result)
    out = {
        "requests": len(reqs),
        "arrivals_per_s": args.rate,
        "mix": "70% LI basins 512..4096 dx10, 30% LI+Exner reaches 1024x256..4096x1024 dx2",
        "simulated_s": 3600.0 * args.sim_hours,
        "scenarios_per_h": round(stats["scenarios_per_h"], 1),
        "p50_s": round(stats["p50_s"], 3),
        "p99_s": round(stats["p99_s"], 3),
        "per_kind_p50_s": {k: round(float(np.percentile(v, 50)), 3) for k, v in kinds.items()},
        "completed": stats["completed"],
        "rejected": stats["rejected"],
        "peak_live": stats["peak_live"],
        "peak_resident_gb": round(stats["peak_resident_gb"], 2),
        "wall_s": round(stats["wall_s"], 1),
        "steps": int(sum(r.steps for r in srv.results)),
    }
    if cap:
        out["capacity"] = cap
    return out


def bench_serve(args) -> int:
    """The serving loop as its own JSON line (serve_figures over ``--requests``)."""
    f = serve_figures(args, args.requests, args.capacity_sample)
    out = {
        "metric": METRIC,
        "value": f["scenarios_per_h"],
        "unit": "scenarios/h",
        "n_gpus": 1,
        "steps": f["steps"],
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"serve_{f['requests']}_poisson_{args.rate}", "mix": f["mix"],
                   "simulated_s": f["simulated_s"]},
        **{k: v for k, v in f.items() if k not in ("scenarios_per_h", "steps", "mix")},
    }
    print(json.dumps(out), flush=True)
    return 0


def default_batch_variants() -> tuple[int, int]:
    """The library's batch defaults: (members without, with the bedload integral)."""
    from rbs.ensemble import Batch

    b = Batch(1)
    try:
        return b.variants()
    finally:
        b.close()


def serve_capacity(reqs, chunk: int, variants=None) -> dict:
    """Capacity of the server against plain per-request launches on the same requests:
    the server takes them as one burst (all arrive at once); plain runs them one after
    another through the reference's tick loop.  Every served result must equal its
    plain result bit for bit."""
    import dataclasses

    from rbs.serve import Server, plain

    burst = [dataclasses.replace(r, arrival_s=0.0) for r in reqs]
    srv = Server(capacity=256, chunk=chunk, keep_results=True, variants=variants)
    st = srv.serve(burst, log=log)
    served = {r.rid: r for r in srv.results}
    t_plain, same = 0.0, True
    for r in reqs:
        res = plain(r)
        t_plain += res.done_s  # a plain run starts at 0: done_s is its wall time
        same = same and res.steps == served[r.rid].steps and np.array_equal(res.h, served[r.rid].h)
    n = len(reqs)
    log(f"[serve] capacity on {n} requests: served burst {st['wall_s']:.1f} s, "
        f"plain {t_plain:.1f} s, identical={same}")
    return {"requests": n, "served_burst_wall_s": round(st["wall_s"], 2),
            "plain_launches_wall_s": round(t_plain, 2),
            "served_scenarios_per_h": round(3600.0 * n / st["wall_s"], 1),
            "plain_launches_scenarios_per_h": round(3600.0 * n / t_plain, 1),
            "speedup_vs_plain_launches": round(t_plain / st["wall_s"], 3),
            "bitwise_identical_to_plain": bool(same)}


def bench_output(args) -> int:
    """The output cadence on the headline basin (4096^2 rain on terrain): 1800 s of
    storm with a frame (depth, u, v) every 300 s plus the baseline, written to a
    result store, three ways -- no output, device frames written by a writer thread,
    and the reference's synchronous read-back per output tick (run.py:389-400)."""
    import shutil
    import tempfile

    from rbs.core.grid import Grid
    from rbs.driver import DeviceRun
    from rbs.io.output import OutputPipeline, run_with_output
    from rbs.io.zarr_store import ZarrWriter, open_store
    from rbs.scheduler import MultiRateScheduler

    n = args.n
    t0 = time.time()
    from rbs import terrain

    d = terrain.rain_on_terrain(n, dx=10.0, seed=args.seed)
    log(f"[output] {n}x{n} inputs in {time.time() - t0:.0f} s")
    rain_sum = float(d["rain"].astype(np.float64).sum())
    end, every = 1800.0, 300.0
    n_frames = int(round(end / every)) + 1
    grid = Grid(ny=n, nx=n, dx=10.0)
    root = Path(tempfile.mkdtemp(prefix="rbs_output_"))

    def fresh():
        from rbs.core.state import State

        st = State.from_bed(d["bed"], d["dx"], depth=0.0, manning=d["manning"])
        st.set_rain_field(d["rain"])
        st.set_infiltration(d["infil"])
        st.set_open_boundaries({e: "open" for e in ("north", "south", "east", "west")})
        st.arm_source_compensation()
        sched = MultiRateScheduler(end_time=end, output_every=every, events=[7200.0])
        run = DeviceRun(st, "local_inertial", sched, alpha=0.7, dt_max=30.0, rain_is_field=True,
                        rain_duration=7200.0, rain_field_sum=rain_sum)
        return st, sched, run

    results = {}
    try:
        # 1. no output
        st, sched, run = fresh()
        st.sync()
        t0 = time.perf_counter()
        steps = sum(run.run_until(seg.t_stop) for seg in sched.segments())
        st.sync()
        results["none"] = time.perf_counter() - t0
        st.close()
        log(f"[output] no output: {steps} steps in {results['none']:.3f} s")
        # 2. device frames, writer thread
        st, sched, run = fresh()
        w = ZarrWriter(root / "async.zarr", grid, n_frames, {"dx": 10.0})
        pipe = OutputPipeline(st, w, slots=3)
        st.sync()
        t0 = time.perf_counter()
        steps_b = run_with_output(run, sched, pipe)
        st.sync()
        stepped = time.perf_counter() - t0
        pipe.close()
        results["async"] = time.perf_counter() - t0
        results["async_stepping"] = stepped
        w.finalize()
        st.close()
        log(f"[output] async frames: stepping done in {stepped:.3f} s, all {pipe.written} "
            f"frames written at {results['async']:.3f} s")
        # 2b. device frames taken and copied, not written (the device path alone)
        class _Discard:
            def append(self, *a):
                pass

        st, sched, run = fresh()
        pipe = OutputPipeline(st, _Discard(), slots=3)
        st.sync()
        t0 = time.perf_counter()
        run_with_output(run, sched, pipe)
        st.sync()
        pipe.close()
        results["frames_only"] = time.perf_counter() - t0
        st.close()
        log(f"[output] frames captured and copied, not written: {results['frames_only']:.3f} s")
        # 3. synchronous read-back per output tick
        st, sched, run = fresh()
        w = ZarrWriter(root / "sync.zarr", grid, n_frames, {"dx": 10.0})
        st.sync()
        t0 = time.perf_counter()
        w.append(0.0, st.h.numpy(), *st.velocities_numpy())
        for seg in sched.segments():
            run.run_until(seg.t_stop)
            if seg.is_output:
                w.append(seg.t_stop, st.h.numpy(), *st.velocities_numpy())
        results["sync"] = time.perf_counter() - t0
        w.finalize()
        st.close()
        log(f"[output] synchronous read-back: {results['sync']:.3f} s")
        a, b = open_store(root / "async.zarr"), open_store(root / "sync.zarr")
        same = all(np.array_equal(a[f].frame(i), b[f].frame(i))
                   for f in ("depth", "u", "v") for i in range(n_frames))
        log(f"[output] async store == synchronous store: {same}")
    finally:
        shutil.rmtree(root, ignore_errors=True)
    frame_mb = 3 * n * n * 4 / 1e6
    out = {
        "metric": METRIC,
        "value": round(n * n * steps_b / results["async"] / 1e9, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": steps_b,
        "warmup": 0,
        "ms_per_step": round(results["async"] * 1e3 / steps_b, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"li_rain_output_{n}", "grid": f"{n}x{n}", "dx_m": 10.0,
                   "frames": n_frames, "frame_fields": "depth, u, v (f32)",
                   "frame_mb": round(frame_mb, 1), "sim_s": end, "output_every_s": every},
        "wall_s": {k: round(v, 3) for k, v in results.items()},
        "output_overhead_ms_per_frame": {
            "async": round((results["async"] - results["none"]) * 1e3 / n_frames, 1),
            "async_on_stepping": round((results["async_stepping"] - results["none"]) * 1e3
                                       / n_frames, 1),
            "frames_only": round((results["frames_only"] - results["none"]) * 1e3 / n_frames, 1),
            "sync": round((results["sync"] - results["none"]) * 1e3 / n_frames, 1)},
        "stores_identical": bool(same),
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_reservoir(args) -> int:
    """A reservoir release on the headline basin (4096^2): the pool is the upper half
    of the grid (8.4 M cells), the outlet a 64 x 64 box; a fixed 500 m^3/s rule every
    300 s through 1800 s of storm, device-driven, against the same run without it."""
    from rbs import terrain
    from rbs.core.state import State
    from rbs.driver import DeviceRun
    from rbs.processes.reservoir import ReservoirOperator, Structure
    from rbs.scheduler import MultiRateScheduler

    n = args.n
    d = terrain.rain_on_terrain(n, dx=10.0, seed=args.seed)
    rain_sum = float(d["rain"].astype(np.float64).sum())
    dam = Structure(name="basin", kind="dam", cells=[(n // 2, 0)], crest_m=0.0,
                    release_rule="fixed", release_m3_s=500.0, pool=(0, 0, n // 2 - 1, n - 1),
                    outlet=(3 * n // 4, n // 2, 3 * n // 4 + 63, n // 2 + 63), interval_s=300.0)
    end = 1800.0

    def run(with_dam: bool):
        st = State.from_bed(d["bed"], d["dx"], depth=0.05, manning=d["manning"])
        st.set_rain_field(d["rain"])
        st.arm_source_compensation()
        op = ReservoirOperator(dam, st) if with_dam else None
        procs = [op.as_slow_process()] if op else []
        sched = MultiRateScheduler(end_time=end, output_every=end, processes=procs)
        dr = DeviceRun(st, "local_inertial", sched, alpha=0.7, dt_max=30.0, rain_is_field=True,
                       rain_duration=7200.0, rain_field_sum=rain_sum)
        release_s = []
        st.sync()
        t0 = time.perf_counter()
        steps = 0
        for seg in sched.segments():
            steps += dr.run_until(seg.t_stop)
            for proc, el in seg.due:
                t1 = time.perf_counter()
                proc.advance(seg.t_stop, el)
                release_s.append(time.perf_counter() - t1)
        st.sync()
        wall = time.perf_counter() - t0
        recs = op.series if op else []
        st.close()
        return wall, steps, release_s, recs

    base, steps0, _, _ = run(False)
    wall, steps1, rel, recs = run(True)
    log(f"[reservoir] without: {steps0} steps {base:.3f} s; with: {steps1} steps {wall:.3f} s; "
        f"releases {[round(1e3 * x, 2) for x in rel]} ms, volumes "
        f"{[round(r['volume_m3']) for r in recs]} m3")
    out = {
        "metric": METRIC,
        "value": round(n * n * steps1 / wall / 1e9, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": steps1, "warmup": 0,
        "ms_per_step": round(wall * 1e3 / steps1, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"li_rain_reservoir_{n}", "grid": f"{n}x{n}",
                   "pool_cells": (n // 2) * n, "outlet_cells": 64 * 64,
                   "activations": len(rel), "interval_s": 300.0, "sim_s": end},
        "release_ms": {"median": round(1e3 * float(np.median(rel)), 3),
                       "max": round(1e3 * max(rel), 3)},
        "wall_s": {"without": round(base, 3), "with": round(wall, 3)},
        "overhead_ms_per_activation": round((wall - base) * 1e3 / max(len(rel), 1), 2),
        "released_m3": [round(r["volume_m3"], 1) for r in recs],
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_reach(args) -> int:
    """The largest served alluvial reach run alone: 4096 x 1024 cells of 2 m, LI with
    the bedload integral, a point-inflow hydrograph on the channel and morphology
    stops every 900 s, 1800 s simulated, device-driven: with the 16-row row blocks
    (which compute every tile), and with the 64 x 8 LDS tile and the 16-row strip
    (which skip dry floodplain tiles); then every variant's step kernel timed on the
    final state.  `value` is the faster skipping run."""
    from rbs import _lib
    from rbs.serve import InputPool, build_state, reach_request

    pool = InputPool(n=1024, seed=args.seed)
    req = reach_request(0, 0.0, pool, np.random.default_rng(args.seed), end_time=1800.0,
                        factor=4.0)
    ny, nx = req.bed.shape
    cells = ny * nx
    count = ctypes.c_int32()
    _lib.call("rbs_li_variant_count", ctypes.byref(count))
    info = []
    for v in range(count.value):
        bx, by, fam, dflt = (ctypes.c_int32() for _ in range(4))
        _lib.call("rbs_li_variant_info", v, ctypes.byref(bx), ctypes.byref(by),
                  ctypes.byref(fam), ctypes.byref(dflt))
        info.append((v, ("lds_tile", "strip", "strip2", "rows")[fam.value], bx.value, by.value))
    tile_v = next(v for v, f, bx, by in info if f == "lds_tile" and (bx, by) == (64, 8))
    strip_v = next(v for v, f, bx, by in info if f == "strip" and by == 16 * 4)
    rows_v = next(v for v, f, bx, by in info if f == "rows" and by == 16)

    runs, sweep = {}, []
    for label, variant in (("rows_16", rows_v), ("lds_tile_64x8", tile_v),
                           ("strip_16", strip_v)):
        st, run, _ = build_state(req, tune_kernel=False)
        _lib.call("rbs_set_li_variant", st._handle, ctypes.c_int32(variant))
        st.sync()
        t0 = time.perf_counter()
        steps = 0
        if args.host_loop:  # the reference's tick loop: one step launch per tick
            from rbs.core import local_inertial as li
            from rbs.processes.inflow import InflowInjector

            inj = InflowInjector(req.inflows, st)
            for tick in run.sched.ticks(lambda: li.compute_dt(st, 0.7, 30.0)):
                inj.apply(st, tick.t0, tick.dt)
                li.step(st, tick.dt, rain=0.0, rain_scale=0.0, t=tick.t0)
                steps += 1
                for proc, el in tick.due:
                    proc.advance(tick.t1, el)
        else:
            for seg in run.sched.segments():
                steps += run.run_until(seg.t_stop)
                for proc, el in seg.due:
                    proc.advance(seg.t_stop, el)
        st.sync()
        wall = time.perf_counter() - t0
        h = st.h.numpy()
        runs[label] = {"variant": variant, "steps": steps, "wall_s": round(wall, 3),
                       "gcell_updates_per_s": round(cells * steps / wall / 1e9, 2),
                       "dry_cell_fraction": round(float((h == 0).mean()), 4)}
        log(f"[reach] {label}: {steps} steps in {wall:.3f} s, "
            f"{cells * steps / wall / 1e9:.1f} Gcell-updates/s, dry {float((h == 0).mean()):.3f}")
        if label == "strip_16":  # every variant's kernel on the final state, no rain
            dt = st.ledger().dt
            for v, fam, bx, by in info:
                _lib.call("rbs_set_li_variant", st._handle, ctypes.c_int32(v))
                avg = ctypes.c_double()
                _lib.call("rbs_time_li_kernel", st._handle, args.kernel_reps, dt, 0.0, 0.0,
                          ctypes.byref(avg))
                sweep.append({"variant": v, "family": fam, "tile": f"{bx}x{by}",
                              "kernel_us": round(avg.value * 1e3, 2)})
                log(f"[reach] variant {v} {fam} {bx}x{by}: {avg.value * 1e3:.1f} us")
        st.close()
    same_steps = len({r["steps"] for r in runs.values()}) == 1
    best = max(("lds_tile_64x8", "strip_16"), key=lambda k: runs[k]["gcell_updates_per_s"])
    out = {
        "metric": METRIC,
        "value": runs[best]["gcell_updates_per_s"],
        "unit": "Gcell-updates/s",
        "n_gpus": 1, "steps": runs[best]["steps"], "warmup": 0,
        "ms_per_step": round(runs[best]["wall_s"] * 1e3 / runs[best]["steps"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"li_exner_reach_{ny}x{nx}", "grid": f"{ny}x{nx}", "dx_m": 2.0,
                   "driver": "host tick loop" if args.host_loop else "device-driven graphs",
                   "sim_s": 1800.0, "sources": "point inflow hydrograph, bedload integral, "
                   "morphology every 900 s, open south edge"},
        "kernel": best,
        "runs": runs,
        "same_steps": same_steps,
        "kernel_sweep_final_state": sweep,
    }
    print(json.dumps(out), flush=True)
    return 0


CHAN_BYTES_PER_CELL = 84  # + w, d, n_ch, the four old components, and the four new


def bench_channels(args) -> int:
    """Reach-scale basin with sub-grid channels: 4096 x 4096 cells of 100 m, rivers
    sized from drained area, rain field (Kahan) + infiltration, open edges; the fused
    channel kernel device-driven from a captured graph."""
    from rbs import _lib, terrain
    from rbs.core.channels import arm_channels
    from rbs.core.state import State
    from rbs.driver import DeviceRun
    from rbs.scheduler import MultiRateScheduler

    n = args.n
    t0 = time.time()
    d = terrain.reach_basin(n, dx=100.0, seed=args.seed)
    log(f"[channels] reach basin {n}x{n} generated in {time.time() - t0:.0f} s, "
        f"{100 * float((d['width'] > 0).mean()):.1f}% channel cells")
    st = State.from_bed(d["bed"], d["dx"], depth=0.0, manning=d["manning"])
    st.set_rain_field(d["rain"])
    st.set_infiltration(d["infil"])
    st.set_open_boundaries({e: "open" for e in ("north", "south", "east", "west")})
    st.arm_source_compensation()
    arm_channels(st, d["width"], d["depth"], d["chan_manning"])
    if args.variant >= 0:
        _lib.call("rbs_set_chan_variant", st._handle, args.variant)
    cells = n * n
    rain_sum = float(d["rain"].astype(np.float64).sum())
    end_time = 21600.0
    sched = MultiRateScheduler(end_time=end_time, output_every=1800.0, events=[7200.0])
    run = DeviceRun(st, "local_inertial", sched, alpha=0.7, dt_max=30.0, rain_is_field=True,
                    rain_duration=7200.0, rain_field_sum=rain_sum)
    t0 = time.time()
    n_spin = 0
    for seg in sched.segments():
        if seg.t_stop > args.spinup + 1e-6:
            break
        n_spin += run.run_until(seg.t_stop)
    log(f"[channels] spin-up to t={st.ledger().t:.0f} s: {n_spin} steps in {time.time() - t0:.1f} s")
    run.run_until(end_time, max_steps=args.warmup)
    st.sync()
    led0 = st.ledger()
    t0 = time.perf_counter()
    taken = run.run_until(end_time, max_steps=args.steps)
    st.sync()
    el = time.perf_counter() - t0
    led1 = st.ledger()
    value = cells * taken / el / 1e9
    log(f"[channels] timed {taken} steps in {el:.3f} s: {value:.2f} Gcell/s; "
        f"t {led0.t:.0f} -> {led1.t:.0f} s, deepest column {led1.speed_max:.2f} m")
    dt_k = float(led1.dt) if led1.dt > 0 else 1.0
    avg = ctypes.c_double()
    _lib.call("rbs_time_li_kernel", st._handle, args.kernel_reps, dt_k, 0.0, 1.0,
              ctypes.byref(avg))
    kern_us = avg.value * 1e3
    achieved = CHAN_BYTES_PER_CELL * cells / (kern_us * 1e-6) / 1e9
    log(f"[channels] fused channel kernel {kern_us:.1f} us -> {achieved:.0f} GB/s "
        f"({100 * achieved / HBM_PEAK_GBS:.1f}% of 8 TB/s)")
    sweep = None
    if args.sweep:
        count = ctypes.c_int32()
        _lib.call("rbs_chan_variant_count", ctypes.byref(count))
        sweep = []
        for v in range(count.value):
            bx, by, dflt = (ctypes.c_int32() for _ in range(3))
            _lib.call("rbs_chan_variant_info", v, ctypes.byref(bx), ctypes.byref(by),
                      ctypes.byref(dflt))
            _lib.call("rbs_set_chan_variant", st._handle, v)
            _lib.call("rbs_time_li_kernel", st._handle, 3, dt_k, 0.0, 1.0, ctypes.byref(avg))
            _lib.call("rbs_time_li_kernel", st._handle, args.kernel_reps, dt_k, 0.0, 1.0,
                      ctypes.byref(avg))
            us = avg.value * 1e3
            gbs = CHAN_BYTES_PER_CELL * cells / (us * 1e-6) / 1e9
            sweep.append({"variant": v, "tile": f"{bx.value}x{by.value}",
                          "default": bool(dflt.value), "kernel_us": round(us, 1),
                          "frac": round(gbs / HBM_PEAK_GBS, 4)})
            log(f"[channels] variant {v} {bx.value}x{by.value}: {us:.1f} us, "
                f"{100 * gbs / HBM_PEAK_GBS:.1f}% of 8 TB/s")
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": taken,
        "warmup": args.warmup,
        "ms_per_step": round(el * 1e3 / max(taken, 1), 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"li_subgrid_channels_{n}", "grid": f"{n}x{n}", "dx_m": 100.0,
                   "scheme": "local_inertial + sub-grid channels",
                   "channel_cells": int((d["width"] > 0).sum()),
                   "sources": "lognormal rain field (Kahan) + infiltration, open edges"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel_us": round(kern_us, 2),
                     "bytes_per_cell": CHAN_BYTES_PER_CELL},
    }
    if sweep:
        out["variant_sweep"] = sweep
    print(json.dumps(out), flush=True)
    return 0


def valley(n: int, dx: float, seed: int):
    """The HLLC valley's bed and channel columns (rbs.terrain.valley)."""
    from rbs import terrain

    v = terrain.valley(n, dx=dx, seed=seed)
    return v["bed"], v["channel_cols"]


def bench_hllc(args) -> int:
    """HLLC valley: inflow hydrograph into the channel, fixed stage downstream."""
    from rbs import _lib
    from rbs.core.state import State
    from rbs.driver import DeviceRun
    from rbs.processes.inflow import Inflow, InflowInjector
    from rbs.scheduler import MultiRateScheduler

    n = args.n if args.n != 4096 else 8192
    dx = 5.0
    t0 = time.time()
    bed, cells = valley(n, dx, 20261204)
    log(f"[hllc] valley {n}x{n} generated in {time.time() - t0:.0f} s")
    # base flow 50 m3/s at normal depth in the channel: h_n = (q n / sqrt(S))^(3/5)
    q = 50.0 / 40.0
    h_n = (q * 0.035 / np.sqrt(5e-4)) ** 0.6
    depth = np.where(bed[:, :] < bed[:, [n // 2]] + h_n, np.maximum(bed[:, [n // 2]] + h_n - bed, 0.0), 0.0)
    st = State.from_bed(bed, dx, depth=depth.astype(np.float32), manning=0.035)
    stage = float(bed[-1, n // 2] + h_n)
    st.set_open_boundaries({"south": "fixed_stage"}, {"south": [(0.0, stage)]})
    hg = [(0.0, 50.0), (3600.0, 800.0), (14400.0, 50.0)]
    InflowInjector([Inflow((0, j), [(t, qq / len(cells)) for t, qq in hg]) for j in cells], st)
    sched = MultiRateScheduler(end_time=14400.0, output_every=600.0, events=[3600.0])
    if args.hllc_variant >= 0:
        _lib.call("rbs_set_hllc_variant", st._handle, args.hllc_variant)
    run = DeviceRun(st, "hllc_fv", sched, alpha=0.45, dt_max=30.0)
    run.run_until(14400.0, max_steps=args.warmup)
    st.sync()
    t0 = time.perf_counter()
    taken = run.run_until(14400.0, max_steps=args.steps)
    st.sync()
    el = time.perf_counter() - t0
    value = n * n * taken / el / 1e9
    led = st.ledger()
    log(f"[hllc] {taken} steps in {el:.2f} s -> {value:.2f} Gcell/s, t={led.t:.1f}")
    dt_k = float(led.dt) if led.dt > 0 else 0.5
    avg = ctypes.c_double()
    _lib.call("rbs_time_hllc_step", st._handle, 10, dt_k, ctypes.byref(avg))
    step_us = avg.value * 1e3
    log(f"[hllc] stage kernels {step_us:.0f} us per step (HIP events)")
    sweep = hllc_sweep(st, dt_k, n * n) if args.sweep else None
    st.close()
    wet_sweep = hllc_wet_sweep(bed, dx) if args.sweep else None
    wet = hllc_wet(bed, dx)
    early = {"steps": taken, "wall_s": round(el, 3), "gcell_s": round(value, 3),
             "t_end_s": round(led.t, 1), "stage_kernels_us_per_step": round(step_us, 1)}
    flood = hllc_flood(args, bed, depth, dx, stage, cells, hg) if args.hllc_hours > 0 else None
    if flood:  # the configuration's own run: the whole flood, every step device-driven
        value, taken, el = flood["gcell_s"], flood["steps"], flood["wall_s"]
    lake = lake_at_rest(bed, dx)
    bytes_cell = 32
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": taken,
        "warmup": args.warmup,
        "ms_per_step": round(el * 1e3 / max(taken, 1), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"hllc_valley_{n}", "grid": f"{n}x{n}", "dx_m": dx,
                   "scheme": "hllc_fv",
                   "variant": args.hllc_variant if args.hllc_variant >= 0 else "default"},
        "roofline": hllc_roofline(n, args.hllc_variant, step_us, bytes_cell),
        "lake_at_rest": lake,
        "wet_lake": wet,
        "early_window": early,
        "flood": flood,
    }
    if sweep:
        out["variant_sweep"] = sweep
        out["variant_sweep_wet"] = wet_sweep
    print(json.dumps(out), flush=True)
    return 0


def hllc_roofline(n: int, variant: int, step_us: float, bytes_cell: int) -> dict:
    """The stage kernels' HBM rate: the measured bytes per step of the library's strips
    on the 8192^2 valley (FETCH_SIZE and WRITE_SIZE, calibrated on a 1 GiB copy) over
    their HIP-event time in the early window.  Cell-updates x bytes per cell would not
    do: blocks that are dry and unchanged write nothing and read only their depths."""
    out = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": None, "traffic": None, "kernel_us": round(step_us, 1),
           "bytes_per_computed_cell": bytes_cell}
    tf = ROOT / "profiles" / "hllc_valley_8192_traffic.json"
    if n == 8192 and variant < 0 and tf.exists():
        traffic = json.loads(tf.read_text())["bytes_per_step"]
        achieved = traffic / (step_us * 1e-6) / 1e9
        out.update(achieved=round(achieved, 1), frac=round(achieved / HBM_PEAK_GBS, 4),
                   traffic=traffic)
    return out


def wet_lake(bed, dx: float):
    """The valley's bed under a lake 1 m above its highest point: every cell wet, every
    face computed (the dry paths never apply)."""
    from rbs.core import hllc
    from rbs.core.state import State

    eta = float(bed.max()) + 1.0
    st = State.from_bed(bed, dx, depth=(eta - bed.astype(np.float64)).astype(np.float32),
                        manning=0.035)
    return st, hllc.compute_dt(st, 0.45, 30.0), eta


def hllc_wet_sweep(bed, dx: float) -> list:
    """Every HLLC variant on the wet lake over the valley's bed."""
    st, dt, eta = wet_lake(bed, dx)
    log(f"[hllc] wet sweep: lake at {eta:.1f} m over the {bed.shape[0]}^2 valley, dt {dt:.3f} s")
    rows = hllc_sweep(st, dt, bed.size)
    st.close()
    return rows


def hllc_wet(bed, dx: float) -> dict:
    """The library's HLLC kernels on the wet lake, where every face is computed.  Both
    fractions: HBM at 32 B per cell, and the FP32 vector rate of the stage kernels'
    FLOPs (SQ_INSTS_VALU_FLOPS_FP32 per launch on this state, x 64 lanes,
    `profiles/hllc_wet_lake_8192_valu.json`) over the step's time, against the
    157.3 TFLOP/s FP32 vector peak."""
    from rbs import _lib

    st, dt, eta = wet_lake(bed, dx)
    avg = ctypes.c_double()
    _lib.call("rbs_time_hllc_step", st._handle, 2, dt, ctypes.byref(avg))
    _lib.call("rbs_time_hllc_step", st._handle, 10, dt, ctypes.byref(avg))
    st.close()
    us = avg.value * 1e3
    cells = bed.size
    achieved = 32 * cells / (us * 1e-6) / 1e9
    out = {"grid": f"{bed.shape[0]}x{bed.shape[1]}", "lake_m": round(eta, 1),
           "step_us": round(us, 1), "gcell_s": round(cells / us / 1e3, 2),
           "hbm": {"achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_cell": 32},
           "valu": None}
    vf = ROOT / "profiles" / "hllc_wet_lake_8192_valu.json"
    if vf.exists() and bed.shape == (8192, 8192):
        v = json.loads(vf.read_text())["kernels"].values()
        flops = 64 * sum(k["fp32_flops_per_wave_inst"] for k in v)
        tflops = flops / (us * 1e-6) / 1e12
        out["valu"] = {"bound": "fp32 vector", "achieved": round(tflops, 2),
                       "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": round(tflops / FP32_PEAK_TFLOPS, 4),
                       "fp32_flops_per_cell_step": round(flops / cells, 1),
                       "valu_insts_per_cell_step": round(64 * sum(k["valu_insts"] for k in v) / cells, 1)}
    log(f"[hllc] wet lake {out['grid']}: {us:.0f} us per step, {out['gcell_s']} Gcell/s, "
        f"HBM {100 * out['hbm']['frac']:.1f}%"
        + (f", FP32 VALU {100 * out['valu']['frac']:.1f}%" if out["valu"] else ""))
    return out


def dry_block_fraction(h, rows: int = 16, cols: int = 58, halo: int = 3) -> float:
    """The share of the default HLLC strip blocks whose depths, three cells out (edges
    clamped), are all exactly zero: the blocks a step's stages leave as they are."""
    ny, nx = h.shape
    wet = np.pad((h != 0.0).astype(np.int64), halo, mode="edge")
    c = np.zeros((ny + 2 * halo + 1, nx + 2 * halo + 1), np.int64)
    c[1:, 1:] = wet.cumsum(0).cumsum(1)
    r0 = np.arange(0, ny, rows)
    c0 = np.arange(0, nx, cols)
    r1 = np.minimum(r0 + rows, ny) + 2 * halo
    c1 = np.minimum(c0 + cols, nx) + 2 * halo
    tot = c[r1][:, c1] - c[r0][:, c1] - c[r1][:, c0] + c[r0][:, c0]
    return float((tot == 0).mean())


def hllc_flood(args, bed, depth, dx: float, stage: float, cells, hg) -> dict:
    """The valley's whole flood, device-driven: the hydrograph's rise to its 1 h peak and
    its recession, timed hour by hour (the wet fraction is read between the legs, outside
    the timing); then the stage kernels timed by HIP events on the flooded state, for
    the library's variant and the 16x16 LDS tiles."""
    from rbs import _lib
    from rbs.core.massbalance import MassLedger
    from rbs.core.state import State
    from rbs.driver import DeviceRun
    from rbs.processes.inflow import Inflow, InflowInjector
    from rbs.scheduler import MultiRateScheduler

    n = bed.shape[0]
    t_end = 3600.0 * args.hllc_hours
    st = State.from_bed(bed, dx, depth=depth.astype(np.float32), manning=0.035)
    st.set_open_boundaries({"south": "fixed_stage"}, {"south": [(0.0, stage)]})
    InflowInjector([Inflow((0, j), [(t, qq / len(cells)) for t, qq in hg]) for j in cells], st)
    if args.hllc_variant >= 0:
        _lib.call("rbs_set_hllc_variant", st._handle, args.hllc_variant)
    sched = MultiRateScheduler(end_time=t_end, output_every=600.0, events=[3600.0])
    run = DeviceRun(st, "hllc_fv", sched, alpha=0.45, dt_max=30.0)
    ledger = MassLedger.from_state(st)
    legs, wet, dry_blocks = [], {}, {}
    for t_stop in sorted({min(3600.0 * h, t_end) for h in range(1, int(np.ceil(args.hllc_hours)) + 1)}):
        t0 = time.perf_counter()
        k = run.run_until(t_stop)
        st.sync()
        legs.append((k, time.perf_counter() - t0))
        h = st.h.numpy()
        key = f"{t_stop / 3600.0:g}h"
        wet[key] = round(float((h > 1e-3).mean()), 5)
        dry_blocks[key] = round(dry_block_fraction(h), 4)
        log(f"[hllc flood] t = {t_stop:.0f} s: {k} steps in {legs[-1][1]:.1f} s, "
            f"wet fraction {wet[key]:.4f}, dry blocks {dry_blocks[key]:.4f}")
    steps = sum(k for k, _ in legs)
    wall = sum(w for _, w in legs)
    led = st.ledger()
    mass = ledger.record(st, led.t).rel_error
    dt_k = float(led.dt) if led.dt > 0 else 0.2
    kern = {}
    for label, v in (("library", -1), ("lds_16x16", 4)):
        if v >= 0:
            _lib.call("rbs_set_hllc_variant", st._handle, v)
        avg = ctypes.c_double()
        _lib.call("rbs_time_hllc_step", st._handle, 10, dt_k, ctypes.byref(avg))
        kern[label] = round(avg.value * 1e3, 1)
    st.close()
    out = {"simulated_s": t_end, "steps": steps, "wall_s": round(wall, 2),
           "gcell_s": round(n * n * steps / wall / 1e9, 3),
           "counting": "every cell of the grid is one update per step; a strip block whose "
                       "stage inputs and outputs are all +0 (dry_block_fraction, the default "
                       "58x16 blocks with their 3-cell halo) writes nothing",
           "wet_fraction": wet, "dry_block_fraction": dry_blocks,
           "stage_kernels_us_per_step_at_end": kern,
           "mass_rel_error": mass}  # the reference's MassLedger measure
    log(f"[hllc flood] {steps} steps over {t_end:.0f} s in {wall:.1f} s -> "
        f"{out['gcell_s']:.2f} Gcell/s; stage kernels at the end {kern}")
    return out


def hllc_sweep(st, dt: float, cells: int) -> list:
    """Mean duration of both stage kernels for every HLLC tile variant (same state)."""
    from rbs import _lib

    count = ctypes.c_int32()
    _lib.call("rbs_hllc_variant_count", ctypes.byref(count))
    rows = []
    for v in range(count.value):
        bx, by, dflt = (ctypes.c_int32() for _ in range(3))
        _lib.call("rbs_hllc_variant_info", v, ctypes.byref(bx), ctypes.byref(by),
                  ctypes.byref(dflt))
        _lib.call("rbs_set_hllc_variant", st._handle, v)
        avg = ctypes.c_double()
        _lib.call("rbs_time_hllc_step", st._handle, 2, dt, ctypes.byref(avg))
        _lib.call("rbs_time_hllc_step", st._handle, 10, dt, ctypes.byref(avg))
        us = avg.value * 1e3
        rows.append({"variant": v, "tile": f"{bx.value}x{by.value}", "default": bool(dflt.value),
                     "step_us": round(us, 1), "gcell_s": round(cells / us / 1e3, 2)})
        log(f"[hllc] variant {v} {bx.value}x{by.value}: {us:.0f} us per step, "
            f"{cells / us / 1e3:.1f} Gcell/s")
    return rows


def lake_at_rest(bed, dx: float, side: int = 1024, steps: int = 1000) -> dict:
    """Flat water 10 m above the lowest bed point of a central crop of the valley,
    closed walls: the largest velocity after ``steps`` HLLC steps (hydrostatic
    balance over the channel banks and the floodplain relief), with the bed as given
    and with the run's datum shift ("auto": the whole metres below the lowest cell,
    rbs.core.datum, as solver/run.py:194 applies it)."""
    from rbs.core import datum, hllc
    from rbs.core.state import State

    n = bed.shape[0]
    r0 = (n - side) // 2
    z_raw = np.ascontiguousarray(bed[r0:r0 + side, r0:r0 + side])
    out = {"grid": f"{side}x{side}", "steps": steps}
    for key, z_ref in (("unshifted", 0.0), ("datum_shifted", datum.resolve_datum("auto", z_raw))):
        z = datum.shift_bed(z_raw, z_ref)
        eta = float(z.min()) + 10.0
        st = State.from_bed(z, dx, depth=np.maximum(eta - z, 0.0).astype(np.float32),
                            manning=0.035)
        for _ in range(steps):
            hllc.step(st, hllc.compute_dt(st, 0.45, 30.0))
        u, v = st.velocities_numpy()
        worst = float(max(np.abs(u).max(), np.abs(v).max()))
        wet = float((st.h.numpy() > 1e-3).mean())
        st.close()
        log(f"[hllc] lake at rest {side}x{side} ({key}, z_ref {z_ref:.0f} m), {steps} steps, "
            f"wet fraction {wet:.2f}: max|u,v| = {worst:.2e} m/s")
        out[key] = {"z_ref_m": z_ref, "wet_fraction": round(wet, 3), "max_velocity_m_s": worst}
    out["max_velocity_m_s"] = out["datum_shifted"]["max_velocity_m_s"]
    out["reference_gate_m_s"] = 1e-4  # validation/test_hllc_2d.py:54
    return out


if __name__ == "__main__":
    sys.exit(main())
