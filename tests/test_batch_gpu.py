"""Batched ensemble: every member's result equals its solo run, bit for bit.

The identity a batched/served ensemble must keep (each result the same alone and in
the mix): members of different sizes, rain rates and stop times advance together,
finish at different times, and a member admitted while others are running (the
continuous-batching case) still matches its solo run.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(members, end):
    from rbs.driver import DeviceRun
    from rbs.ensemble import make_member_state
    from rbs.scheduler import MultiRateScheduler

    out = []
    for m in members:
        st = make_member_state(m)
        sched = MultiRateScheduler(end_time=end, output_every=end, events=[0.75 * end])
        run = DeviceRun(st, "local_inertial", sched, alpha=0.7, dt_max=30.0, rain_is_field=True,
                        rain_duration=0.75 * end,
                        rain_field_sum=float(m.rain.astype(np.float64).sum()))
        out.append(run)
    return out


def test_batch_members_equal_solo_runs():
    from rbs.ensemble import Batch, ensemble_inputs

    members = ensemble_inputs(6, seed=7, n_min=64, n_max=300)
    ends = [600.0, 900.0, 600.0, 1200.0, 300.0, 900.0]
    solo = [_setup([m], e)[0] for m, e in zip(members, ends)]
    for run, e in zip(solo, ends):
        run.run_until(e)
    batched = [_setup([m], e)[0] for m, e in zip(members, ends)]
    b = Batch(8)
    slots = [b.add(r, e) for r, e in zip(batched[:4], ends[:4])]
    b.advance(40)  # the last two join while the first four are running
    slots += [b.add(r, e) for r, e in zip(batched[4:], ends[4:])]
    b.run_to_completion(32)
    for s in slots:
        b.remove(s)
    for a, c in zip(solo, batched):
        la, lc = a.state.ledger(), c.state.ledger()
        assert la.steps == lc.steps and la.t == lc.t
        assert np.array_equal(a.state.h.numpy(), c.state.h.numpy())
        assert np.array_equal(a.state.qx.numpy(), c.state.qx.numpy())
        assert np.array_equal(a.state.qy.numpy(), c.state.qy.numpy())
        assert la.rain_in == lc.rain_in and la.edge_out == lc.edge_out
    b.close()


def test_row_block_batches_equal_solo_runs():
    """The row-block kernel (eight strips per block, rows stored through LDS, an idle
    wave past each member's east edge) in a batch: bitwise its solo runs, ledger too."""
    import ctypes

    from rbs import _lib
    from rbs.autotune import variant_table
    from rbs.ensemble import Batch, ensemble_inputs

    rows = [r["variant"] for r in variant_table() if r["family"] == "rows"]
    assert rows
    members = ensemble_inputs(5, seed=11, n_min=100, n_max=700)
    ends = [600.0, 300.0, 900.0, 600.0, 450.0]
    for v in rows:
        solo = [_setup([m], e)[0] for m, e in zip(members, ends)]
        for run, e in zip(solo, ends):
            _lib.call("rbs_set_li_variant", run.state._handle, ctypes.c_int32(v))
            run.run_until(e)
        batched = [_setup([m], e)[0] for m, e in zip(members, ends)]
        b = Batch(8)
        b.set_variants(v, b.variants()[1])
        slots = [b.add(r, e) for r, e in zip(batched[:3], ends[:3])]
        b.advance(24)
        slots += [b.add(r, e) for r, e in zip(batched[3:], ends[3:])]
        b.run_to_completion(32)
        for s in slots:
            b.remove(s)
        for a, c in zip(solo, batched):
            la, lc = a.state.ledger(), c.state.ledger()
            assert la.steps == lc.steps and la.t == lc.t, v
            assert np.array_equal(a.state.h.numpy(), c.state.h.numpy()), v
            assert np.array_equal(a.state.qx.numpy(), c.state.qx.numpy()), v
            assert np.array_equal(a.state.qy.numpy(), c.state.qy.numpy()), v
            assert la.rain_in == lc.rain_in and la.edge_out == lc.edge_out, v
            assert la.volume == lc.volume, v
        b.close()
