"""Dry tiles: the LDS-tile and 60-column strip kernels skip a tile whose depths, one
cell out, are all +0 at a step without rain, when a per-tile flag says that both
ping-pong buffers and the Kahan debts of its cells hold +0 -- everything the step
would write is then already in memory.  The skipping kernels must write the bits of
the row-block kernels (which compute every tile) on runs that are mostly dry: an
alluvial reach with a dry floodplain, point inflows, the bedload integral and
morphology stops; a basin whose rain stops and whose infiltration takes cells back to
0; and runs with an upload and a change of tiling in mid-run, which clear the flags.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LEDGER = ("steps", "t", "volume", "rain_in", "point_in", "infil_out", "edge_out")


def _families():
    """The reference kernel (a row block: computes every tile) and the skipping ones."""
    from rbs.autotune import variant_table

    table = variant_table()
    rows = [r["variant"] for r in table if r["family"] == "rows"]
    skipping = [r["variant"] for r in table if r["family"] in ("lds_tile", "strip")]
    assert rows and skipping
    return rows[0], skipping


def _run(req, variant, poke=None):
    from rbs import _lib
    from rbs.serve import build_state

    st, run, _ = build_state(req, tune_kernel=False)
    _lib.call("rbs_set_li_variant", st._handle, ctypes.c_int32(variant))
    for k, seg in enumerate(run.sched.segments()):
        run.run_until(seg.t_stop)
        for proc, el in seg.due:
            proc.advance(seg.t_stop, el)
        if poke is not None:
            poke(st, k)
    out = {"h": st.h.numpy(), "qx": st.qx.numpy(), "qy": st.qy.numpy(), "z": st.z.numpy()}
    if st.sediment is not None:
        sed = st.sediment
        out.update(qs_int_x=sed.qs_int_x.numpy(), qs_comp_x=sed.qs_comp_x.numpy(),
                   qs_int_y=sed.qs_int_y.numpy(), qs_comp_y=sed.qs_comp_y.numpy())
    led = st.ledger()
    out.update({f"ledger.{k}": getattr(led, k) for k in LEDGER})
    st.close()
    return out


def _same(a, b, label):
    """Fields bit for bit; the ledger's clock exactly, and its sums of per-tile
    partials to the rounding of their order (tilings differ in how many partials a
    sum has)."""
    assert a.keys() == b.keys()
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert np.array_equal(a[k], b[k]), f"{label}: {k} differs"
        elif k in ("ledger.steps", "ledger.t"):
            assert a[k] == b[k], f"{label}: {k} {a[k]!r} != {b[k]!r}"
        else:
            assert a[k] == pytest.approx(b[k], rel=1e-12, abs=1e-9), f"{label}: {k}"


@pytest.fixture(scope="module")
def pool():
    from rbs import _lib
    from rbs.serve import InputPool

    _lib.load()
    return InputPool(n=512, seed=3)


def test_a_mostly_dry_reach_skips_tiles_and_keeps_every_bit(pool):
    from rbs.serve import reach_request

    req = reach_request(0, 0.0, pool, np.random.default_rng(4), end_time=1800.0, scale=0.25)
    req.interval_s = 600.0  # two morphology stops inside the run
    ref, skipping = _families()
    want = _run(req, ref)
    assert (want["h"] == 0).mean() > 0.6  # most of the floodplain never wets
    assert not np.array_equal(want["z"], req.bed)  # the stops moved the bed
    for v in skipping:
        _same(want, _run(req, v), f"variant {v}")


def test_a_basin_drying_after_its_rain_keeps_every_bit(pool):
    from rbs.serve import basin_request

    req = basin_request(0, 0.0, pool, np.random.default_rng(6), n_min=150, n_max=200,
                        end_time=3600.0)
    req.infil = req.infil * np.float32(20.0)  # strong enough to dry hillslopes to 0
    ref, skipping = _families()
    want = _run(req, ref)
    assert (want["h"] == 0).any()
    for v in skipping:
        _same(want, _run(req, v), f"variant {v}")


def test_uploads_and_tiling_changes_in_mid_run_clear_the_flags(pool):
    """Water poured onto the dry floodplain by an upload after the first segment, and
    the kernel switched to another strip shape after the second: both must be seen by
    tiles that were skipping."""
    from rbs import _lib
    from rbs.serve import reach_request

    req = reach_request(1, 0.0, pool, np.random.default_rng(8), end_time=1800.0, scale=0.25)
    req.interval_s = 600.0
    ref, skipping = _families()
    ny, nx = req.bed.shape

    def poke(switch_to):
        def f(st, k):
            if k == 0:
                h = st.h.numpy()
                assert not h[ny // 3:ny // 3 + 5, 2:8].any()  # dry there before the pour
                h[ny // 3:ny // 3 + 5, 2:8] += np.float32(0.4)
                st.h.assign(h)
            elif k == 1 and switch_to is not None:
                _lib.call("rbs_set_li_variant", st._handle, ctypes.c_int32(switch_to))
        return f

    want = _run(req, ref, poke(None))
    assert want["h"][:, :10].any()  # the poured water is still on the floodplain
    for a, b in zip(skipping, skipping[1:] + skipping[:1]):
        _same(want, _run(req, a, poke(b)), f"variant {a} -> {b}")


def test_a_state_moved_between_kernels_by_a_batch_keeps_every_bit(pool):
    """A reach stepped alone on the 64 x 8 LDS tile (its flags set for that tiling),
    then in a batch, which moves it to the bedload strips and clears its flags in the
    batch's stream order, then alone again on the batch's kernel: the bits of a
    reach that stops at the same times on the row blocks."""
    from rbs import _lib
    from rbs.ensemble import Batch
    from rbs.serve import build_state, reach_request

    req = reach_request(2, 0.0, pool, np.random.default_rng(10), end_time=1800.0, scale=0.25)
    req.interval_s = 1.0e6  # the bedload integral without activations
    ref, _ = _families()

    def states(variant):
        st, run, _ = build_state(req, tune_kernel=False)
        _lib.call("rbs_set_li_variant", st._handle, ctypes.c_int32(variant))
        return st, run

    st_ref, run_ref = states(ref)
    for t in (600.0, 1200.0, 1800.0):
        run_ref.run_until(t)
    st, run = states(_tile_variants()[0])
    run.run_until(600.0)
    b = Batch(4)
    slot = b.add(run, 1200.0)
    b.run_to_completion(16)
    b.remove(slot)
    b.close()
    run.run_until(1800.0)
    for name in ("h", "qx", "qy"):
        assert np.array_equal(getattr(st, name).numpy(), getattr(st_ref, name).numpy()), name
    for name in ("qs_int_x", "qs_comp_x", "qs_int_y", "qs_comp_y"):
        assert np.array_equal(getattr(st.sediment, name).numpy(),
                              getattr(st_ref.sediment, name).numpy()), name
    assert st.ledger().steps == st_ref.ledger().steps and st.ledger().t == st_ref.ledger().t
    assert (st.h.numpy() == 0).mean() > 0.6
    st.close()
    st_ref.close()


def _tile_variants():
    from rbs.autotune import variant_table

    return [r["variant"] for r in variant_table()
            if r["family"] == "lds_tile" and r["tile"] == (64, 8)]
