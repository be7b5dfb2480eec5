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
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--n", type=int, default=4096, help="grid side (cells)")
    ap.add_argument("--spinup", type=float, default=3600.0, help="simulated s before timing")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--variant", type=int, default=-1, help="LI tile variant (-1: tuned)")
    ap.add_argument("--kernel-reps", type=int, default=50)
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(a scenario shards by row strips; see DESIGN.md)", file=sys.stderr)
        return 2

    from rbs import _lib
    from rbs.driver import DeviceRun
    from rbs.scheduler import MultiRateScheduler

    _lib.load()
    st, d = build_state(args.n, args.seed)
    if args.variant >= 0:
        _lib.call("rbs_set_li_variant", st._handle, args.variant)
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
    traffic = None
    tf = ROOT / "profiles" / "li_rain_4096_traffic.json"
    if tf.exists() and args.n == 4096:
        traffic = json.loads(tf.read_text()).get("bytes_per_launch")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(st, d, dt_k, min(16, os.cpu_count() or 1), args.cpu_budget)

    mass = None
    try:
        inflow = led1.rain_in
        outflow = led1.infil_out + led1.edge_out
        mass = abs(inflow - outflow - (led1.volume - 0.0)) / max(inflow + outflow, 1e-12)
    except Exception:
        pass

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
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
