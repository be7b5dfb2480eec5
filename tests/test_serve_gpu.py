"""Serving: continuous batching keeps every result identical to the request alone.

Requests of both kinds (rain-on-terrain basins with infiltration; alluvial reaches
with point-source inflow and Exner morphology stopping at activations) arrive
while others run; each served result must equal its solo run bit for bit.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_served_results_equal_solo_results():
    from rbs.serve import InputPool, Server, make_requests, solo

    pool = InputPool(n=512, seed=11)
    reqs = make_requests(8, rate_per_s=40.0, seed=11, pool=pool, end_time=600.0,
                         basin_min=96, basin_max=300, reach_scale=0.125)
    for r in reqs:
        if r.kind == "reach":
            r.interval_s = 200.0  # several morphology stops inside 600 s
    assert {r.kind for r in reqs} == {"basin", "reach"}
    srv = Server(capacity=16, chunk=16, keep_results=True)
    stats = srv.serve(reqs)
    assert stats["completed"] == len(reqs) and stats["rejected"] == 0
    assert stats["peak_live"] >= 2  # requests overlapped on the device
    by_id = {r.rid: r for r in srv.results}
    for req in reqs:
        alone = solo(req)
        got = by_id[req.rid]
        assert got.steps == alone.steps
        assert np.array_equal(got.h, alone.h), f"request {req.rid} ({req.kind}) differs"


def test_admission_refuses_what_cannot_fit():
    from rbs.serve import InputPool, Server, make_requests

    pool = InputPool(n=256, seed=3)
    reqs = make_requests(3, rate_per_s=100.0, seed=3, pool=pool, end_time=60.0,
                         basin_min=64, basin_max=200)
    for r in reqs:
        r.kind = "basin"
        r.d50 = None
        r.inflows = []
    srv = Server(capacity=4, hbm_budget=reqs[0].footprint_bytes() - 1)
    stats = srv.serve(reqs)
    assert stats["rejected"] >= 1


def test_plain_launches_equal_device_driven_solo():
    """The reference-style tick loop (compute_dt, host fill_span, sources, one step per
    tick, morphology at its activations) and the device-driven solo run agree bit for
    bit, for a basin and for a reach with point sources and Exner stops."""
    from rbs.serve import InputPool, make_requests, plain, solo

    pool = InputPool(n=256, seed=5)
    reqs = make_requests(6, rate_per_s=10.0, seed=5, pool=pool, end_time=400.0,
                         basin_min=64, basin_max=160, reach_scale=0.1)
    kinds = set()
    for r in reqs:
        if r.kind == "reach":
            r.interval_s = 100.0
        if r.kind in kinds:
            continue
        kinds.add(r.kind)
        a, b = plain(r), solo(r)
        assert a.steps == b.steps
        assert np.array_equal(a.h, b.h), f"request {r.rid} ({r.kind}) differs"
    assert kinds == {"basin", "reach"}


def test_burst_builds_overlap_stepping_and_results_equal_solo():
    """Every request arrives at once: states are built on the preparation thread while
    the batch captures and steps on the main thread; each result equals its solo run."""
    import dataclasses

    from rbs.serve import InputPool, Server, make_requests, solo

    pool = InputPool(n=512, seed=21)
    reqs = make_requests(12, rate_per_s=10.0, seed=21, pool=pool, end_time=300.0,
                         basin_min=96, basin_max=320, reach_scale=0.125)
    reqs = [dataclasses.replace(r, arrival_s=0.0) for r in reqs]
    for r in reqs:
        if r.kind == "reach":
            r.interval_s = 100.0
    srv = Server(capacity=16, chunk=16, keep_results=True, prep_threads=2)
    stats = srv.serve(reqs)
    assert stats["completed"] == len(reqs) and stats["rejected"] == 0
    assert stats["batch_rebuilds"] >= 2
    by_id = {r.rid: r for r in srv.results}
    for req in reqs:
        alone = solo(req)
        assert by_id[req.rid].steps == alone.steps
        assert np.array_equal(by_id[req.rid].h, alone.h), f"request {req.rid} ({req.kind})"


def test_narrow_wide_and_bedload_segments_equal_solo_results():
    """A batch holding all three tile segments at once: narrow basins on the strips, a
    basin at least 2048 columns wide on the row blocks, and reaches with the bedload
    integral, each segment on its own graph branch; every result equals its solo run."""
    from rbs.serve import InputPool, Server, basin_request, reach_request, solo

    pool = InputPool(n=2112, seed=5)
    rng = np.random.default_rng(5)
    # the small requests run long enough to still be live while the wide one is built
    reqs = [basin_request(0, 0.0, pool, rng, n_min=96, n_max=200, end_time=3600.0),
            basin_request(1, 0.0, pool, rng, n_min=2050, n_max=2100, end_time=300.0),
            reach_request(2, 0.0, pool, rng, end_time=3600.0, scale=0.125),
            basin_request(3, 0.01, pool, rng, n_min=96, n_max=200, end_time=3600.0)]
    reqs[2].interval_s = 600.0
    assert reqs[1].bed.shape[1] >= 2048 > max(r.bed.shape[1] for r in reqs if r.rid != 1)
    srv = Server(capacity=8, chunk=16, keep_results=True)
    stats = srv.serve(reqs)
    assert stats["completed"] == len(reqs) and stats["peak_live"] >= 3
    by_id = {r.rid: r for r in srv.results}
    for req in reqs:
        alone = solo(req)
        got = by_id[req.rid]
        assert got.steps == alone.steps
        assert np.array_equal(got.h, alone.h), f"request {req.rid} ({req.kind}) differs"


def test_admission_footprint_is_what_the_library_allocates():
    """rbs_state_footprint's field bytes equal rbs_state_bytes of a fresh state of that
    shape and arming, for every arming a request (or a HLLC or channel run) uses."""
    import ctypes

    from rbs import _lib
    from rbs.core.state import State

    _lib.load()
    for shape in ((97, 131), (512, 2050), (300, 64)):
        for arm in (0, _lib.ARM_COMP | _lib.ARM_RAIN | _lib.ARM_INFIL, _lib.ARM_SED,
                    _lib.ARM_HLLC, _lib.ARM_CHAN, _lib.ARM_COMP | _lib.ARM_RAIN | _lib.ARM_SED):
            fields, total = ctypes.c_uint64(), ctypes.c_uint64()
            _lib.call("rbs_state_footprint", shape[0], shape[1], arm, ctypes.byref(fields),
                      ctypes.byref(total))
            st = State.from_bed(np.zeros(shape, np.float32), 10.0)
            if arm:
                st._arm(arm)
            assert st.device_bytes() == fields.value, (shape, arm)
            assert total.value > fields.value
            st.close()
