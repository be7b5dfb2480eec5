"""The batched-ensemble and serving workloads at their own sizes.

* The ensemble: all 256 scenarios (sides log-uniform 256..2048, dx 10 m, each on its
  own fBm DEM with a lognormal rain field, open edges, 1 h simulated), advanced in one
  batch, each equal bit for bit to the same scenario run alone; the two smallest also
  against the oracle's tick loop (compute_dt, fill_span, step).
* Serving: a mix holding a 4096^2 rain-on-terrain basin and a 4096x1024 alluvial reach
  (LI + Exner, point-source hydrograph, morphology stops) beside smaller requests,
  arriving while the others run; each result equal to its solo run.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

END = 3600.0


def _member_run(m):
    from rbs.core.state import State
    from rbs.driver import DeviceRun
    from rbs.scheduler import MultiRateScheduler

    st = State.from_bed(m.bed, m.dx, depth=0.0, manning=m.manning)
    st.set_rain_field(m.rain)
    st.set_open_boundaries({e: "open" for e in ("north", "south", "east", "west")})
    st.arm_source_compensation()
    sched = MultiRateScheduler(end_time=END, output_every=END)
    return DeviceRun(st, "local_inertial", sched, alpha=0.7, dt_max=30.0, rain_is_field=True,
                     rain_duration=END, rain_field_sum=float(m.rain.astype(np.float64).sum()))


def test_ensemble_of_256_batched_equals_each_scenario_alone(orc):
    from rbs import _lib
    from rbs.ensemble import Batch, ensemble_inputs
    from rbs.scheduler import MultiRateScheduler

    _lib.load()
    members = ensemble_inputs(256, seed=20261205)
    sides = sorted(m.bed.shape[0] for m in members)
    assert len(members) == 256 and sides[0] >= 256 and sides[-1] <= 2048
    print(f"\n[ensemble] 256 scenarios, sides {sides[0]}..{sides[-1]}, "
          f"{sum(m.bed.size for m in members) / 1e6:.0f} M cells", flush=True)
    alone, steps_alone = [], []
    for m in members:
        run = _member_run(m)
        steps_alone.append(run.run_until(END))
        alone.append(run.state.h.numpy())
        run.state.close()
    runs = [_member_run(m) for m in members]
    b = Batch(len(runs))
    slots = [b.add(r, END) for r in runs]
    b.run_to_completion(64)
    for sl in slots:
        b.remove(sl)
    differ = [k for k, r in enumerate(runs) if not np.array_equal(r.state.h.numpy(), alone[k])]
    steps = [int(r.state.ledger().steps) for r in runs]
    print(f"[ensemble] batched: {sum(steps)} steps, {len(differ)} members differ from alone",
          flush=True)
    assert not differ
    assert steps == steps_alone
    # the two smallest scenarios against the oracle's tick loop over the same hour
    orc.set_threads(8)
    for k in sorted(range(256), key=lambda k: members[k].bed.size)[:2]:
        m = members[k]
        o = orc.State(m.bed, m.dx, depth=0.0, manning=m.manning)
        o.set_rain_field(m.rain)
        o.set_open_boundaries({e: "open" for e in ("north", "south", "east", "west")})
        o.arm_source_compensation()
        sched = MultiRateScheduler(end_time=END, output_every=END)
        n = 0
        for tick in sched.ticks(lambda: orc.li_compute_dt(o, 0.7, 30.0)):
            orc.li_step(o, tick.dt, rain_scale=1.0)
            n += 1
        print(f"[ensemble] member {k} ({m.bed.shape[0]}^2) against the oracle: {n} steps, "
              f"max|dh| {float(np.abs(runs[k].state.h.numpy() - o.h).max()):.1e}", flush=True)
        assert n == steps[k]
        assert np.array_equal(runs[k].state.h.numpy(), o.h)
    for r in runs:
        r.state.close()
    b.close()


def test_served_mix_at_full_size_equals_each_request_alone():
    from rbs import _lib
    from rbs.serve import InputPool, Server, basin_request, reach_request, solo

    _lib.load()
    pool = InputPool(n=4096, seed=20261206)
    rng = np.random.default_rng(20261206)
    reqs = [basin_request(0, 0.0, pool, rng, n_min=4096, n_max=4096, end_time=END),
            reach_request(1, 0.05, pool, rng, end_time=END, factor=4.0),
            basin_request(2, 0.1, pool, rng, n_min=512, n_max=1024, end_time=END),
            reach_request(3, 0.15, pool, rng, end_time=END),
            basin_request(4, 0.2, pool, rng, n_min=512, n_max=2048, end_time=END)]
    assert reqs[0].bed.shape == (4096, 4096) and reqs[1].bed.shape == (4096, 1024)
    srv = Server(capacity=16, chunk=32, keep_results=True)
    stats = srv.serve(reqs)
    print(f"\n[serve full size] {stats['completed']} requests in {stats['wall_s']:.1f} s, "
          f"peak live {stats['peak_live']}, p50 {stats['p50_s']:.2f} s, p99 {stats['p99_s']:.2f} s",
          flush=True)
    assert stats["completed"] == len(reqs) and stats["rejected"] == 0
    assert stats["peak_live"] >= 2
    by_id = {r.rid: r for r in srv.results}
    for req in reqs:
        alone = solo(req)
        got = by_id[req.rid]
        assert got.steps == alone.steps, req.rid
        assert np.array_equal(got.h, alone.h), f"request {req.rid} ({req.kind}) differs"
