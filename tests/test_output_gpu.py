"""Output frames on the HIP path: the frame kernel against the reference's read-back
arithmetic (state.py:232-256, zarr_writer.py:147) on downloaded fields, and the
asynchronous output loop against the reference's synchronous one, store for store."""

from __future__ import annotations

import numpy as np
import pytest

from test_output_store import velocities

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rbs():
    from rbs import _lib
    from rbs.core import hllc, local_inertial, sediment, state
    from rbs.io import frames, output, zarr_store
    from rbs.processes import morphology

    _lib.load()
    return type("R", (), dict(lib=_lib, li=local_inertial, hllc=hllc, State=state.State,
                              sed=sediment, morph=morphology, frames=frames, out=output,
                              store=zarr_store))


def _terrain(ny, nx, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:ny, 0:nx]
    return (0.01 * (nx - xx) + 0.005 * yy + 0.3 * rng.uniform(0, 1, (ny, nx))).astype(np.float32)


def test_li_frame_is_the_reference_read_back(rbs):
    z = _terrain(70, 130, 1)
    st = rbs.State.from_bed(z, 5.0, depth=0.0, manning=0.03)
    st.set_open_boundaries({"east": "open"})
    for _ in range(120):
        rbs.li.step(st, rbs.li.compute_dt(st, 0.7, 2.0), rain=2e-4)
    q = rbs.frames.FrameQueue(st, ("depth", "u", "v"), slots=2)
    try:
        slot = q.capture()
        f = q.wait(slot)
        h, qx, qy = st.h.numpy(), st.qx.numpy(), st.qy.numpy()
        u, v = velocities(h, qx, qy)
        assert np.array_equal(f.depth, h)
        assert np.array_equal(f.u, u) and np.array_equal(f.v, v)
        assert np.abs(u).max() > 0 and (h < 1e-3).any()  # both wet and dry cells
        assert f.time == st.ledger().t and f.bed_change is None
        q.release(slot)
    finally:
        q.close()


def test_hllc_frame_uses_the_momentum(rbs):
    z = _terrain(40, 60, 2)
    st = rbs.State.from_bed(z, 2.0, depth=np.where(np.arange(60) < 30, 1.0, 0.0)[None, :]
                            * np.ones((40, 1), np.float32), manning=0.02)
    t = 0.0
    for _ in range(40):
        dt = rbs.hllc.compute_dt(st, 0.45, 1.0)
        rbs.hllc.step(st, dt, t=t)
        t += dt
    q = rbs.frames.FrameQueue(st, ("depth", "u", "v"), slots=1)
    try:
        f = q.wait(q.capture())
        h = st.h.numpy()
        u, v = velocities(h, None, None, st.hu.numpy(), st.hv.numpy())
        assert np.array_equal(f.depth, h) and np.array_equal(f.u, u) and np.array_equal(f.v, v)
        assert np.abs(u).max() > 0
    finally:
        q.close()


def test_bed_change_frame_is_the_float64_change_narrowed(rbs):
    z = _terrain(32, 96, 3)
    st = rbs.State.from_bed(z, 5.0, depth=0.4, manning=0.03)
    rbs.sed.arm_sediment(st, np.full(z.shape, 1e-3, np.float32), 0.4)
    for _ in range(80):
        rbs.li.step(st, rbs.li.compute_dt(st, 0.7, 2.0))
    rbs.morph.MorphologyProcess(st, 100.0).advance(100.0, 100.0)
    q = rbs.frames.FrameQueue(st, ("depth", "bed_change"), slots=1)
    try:
        f = q.wait(q.capture())
        dz = st.sediment.dz_cum.numpy()
        assert np.array_equal(f.bed_change, dz.astype(np.float32))
        assert np.abs(dz).max() > 0 and f.u is None
    finally:
        q.close()


def test_held_slots_refuse_a_capture_until_released(rbs):
    st = rbs.State.from_bed(np.zeros((8, 8), np.float32), 1.0, depth=0.1)
    q = rbs.frames.FrameQueue(st, ("depth",), slots=1)
    try:
        slot = q.capture()
        with pytest.raises(Exception, match="slots are held"):
            q.capture()
        q.release(slot)
        q.release(q.capture())
        with pytest.raises(ValueError):
            rbs.frames.FrameQueue(st, ("u",))  # u without v
        with pytest.raises(Exception, match="RBS_ARM_SED"):
            rbs.frames.FrameQueue(st, ("bed_change",))
    finally:
        q.close()


def test_async_output_equals_the_synchronous_reference_loop(rbs, tmp_path):
    """run.py:347-400 (synchronous read-back per output tick) against
    run_with_output (device frames written by a writer thread): the same stores."""
    from rbs.core.grid import Grid
    from rbs.driver import DeviceRun
    from rbs.scheduler import MultiRateScheduler

    z = _terrain(96, 160, 4)
    rain = 60.0 / 1000 / 3600
    end, every = 1200.0, 150.0
    n_frames = int(round(end / every)) + 1
    grid = Grid(ny=96, nx=160, dx=10.0)

    ref = rbs.State.from_bed(z, 10.0, depth=0.0, manning=0.035)
    ref.set_open_boundaries({"east": "open", "south": "open"})
    w_ref = rbs.store.ZarrWriter(tmp_path / "ref.zarr", grid, n_frames, {"dx": 10.0})
    w_ref.append(0.0, ref.h.numpy(), *ref.velocities_numpy())
    sched = MultiRateScheduler(end_time=end, output_every=every, events=[900.0])
    for tick in sched.ticks(lambda: rbs.li.compute_dt(ref, 0.7, 10.0)):
        rbs.li.step(ref, tick.dt, rain=rain if tick.t0 < 900.0 - 1e-6 else 0.0, t=tick.t0)
        if tick.is_output:
            w_ref.append(tick.t1, ref.h.numpy(), *ref.velocities_numpy())
    w_ref.finalize()

    dev = rbs.State.from_bed(z, 10.0, depth=0.0, manning=0.035)
    dev.set_open_boundaries({"east": "open", "south": "open"})
    w_dev = rbs.store.ZarrWriter(tmp_path / "dev.zarr", grid, n_frames, {"dx": 10.0})
    run = DeviceRun(dev, "local_inertial", sched, alpha=0.7, dt_max=10.0, rain_m_s=rain,
                    rain_duration=900.0)
    pipe = rbs.out.OutputPipeline(dev, w_dev, slots=2)
    try:
        rbs.out.run_with_output(run, sched, pipe)
    finally:
        pipe.close()
    w_dev.finalize()
    assert pipe.written == n_frames

    a, b = rbs.store.open_store(tmp_path / "ref.zarr"), rbs.store.open_store(tmp_path / "dev.zarr")
    assert np.array_equal(a["time"].read(), b["time"].read())
    for name in ("depth", "u", "v"):
        for i in range(n_frames):
            assert np.array_equal(a[name].frame(i), b[name].frame(i)), (name, i)
    assert float(a["depth"].frame(n_frames - 1).max()) > 0.01
