"""The result store and the viewer tiles on the CPU (reference solver/io/test_zarr_writer.py,
solver/io/test_viewer_export.py), with frames from oracle runs.

zarr and xarray are not installed here, so the store is checked against the Zarr v3
format itself (metadata documents, chunk keys, chunk bytes) through the library's own
reader; that zarr-python opens it is unpinned.
"""

from __future__ import annotations

import gzip
import json

import numpy as np
import pytest

from rbs.core.channels import eta_from_h
from rbs.core.grid import H_DRY, Grid
from rbs.io import viewer_export
from rbs.io.viewer_export import export_frames, render_eta
from rbs.io.zarr_store import ZarrWriter, open_store
from rbs.scheduler import MultiRateScheduler


def _grid():
    return Grid(ny=3, nx=4, dx=10.0)


def _frame(g, fill):
    return np.full(g.shape, fill, np.float32)


# --- the store (test_zarr_writer.py) -----------------------------------------------
def test_a_store_without_morphology_has_no_bed_change_array(tmp_path):
    g = _grid()
    w = ZarrWriter(tmp_path / "r.zarr", g, 2, {"dx": g.dx})
    w.write_bed(_frame(g, 100.0))
    for i in range(2):
        w.append(float(i), _frame(g, 0.1), _frame(g, 0.0), _frame(g, 0.0))
    w.finalize()
    st = open_store(tmp_path / "r.zarr")
    assert "bed_change" not in st
    assert set(st.arrays) - {"time"} == {"depth", "u", "v", "bed"}


def test_bed_change_rides_the_time_axis(tmp_path):
    g = _grid()
    w = ZarrWriter(tmp_path / "r.zarr", g, 3, {"dx": g.dx}, bed_change=True)
    w.write_bed(_frame(g, 100.0))
    for i in range(3):
        w.append(float(i), _frame(g, 0.1), _frame(g, 0.0), _frame(g, 0.0), _frame(g, -0.01 * i))
    w.finalize()
    bc = open_store(tmp_path / "r.zarr")["bed_change"]
    assert bc.dims == ("time", "y", "x")
    assert bc.shape == (3, *g.shape) and bc.dtype == np.float32
    assert bc.chunks == (1, *g.shape)
    assert float(bc.frame(2).min()) == pytest.approx(-0.02)


def test_a_missing_or_unexpected_bed_change_frame_is_an_error_both_ways(tmp_path):
    g = _grid()
    morphing = ZarrWriter(tmp_path / "a.zarr", g, 1, {}, bed_change=True)
    with pytest.raises(ValueError, match="bed did not move"):
        morphing.append(0.0, _frame(g, 0.1), _frame(g, 0.0), _frame(g, 0.0))
    still = ZarrWriter(tmp_path / "b.zarr", g, 1, {})
    with pytest.raises(ValueError, match="exactly when the store carries it"):
        still.append(0.0, _frame(g, 0.1), _frame(g, 0.0), _frame(g, 0.0), _frame(g, 0.0))
    with pytest.raises(IndexError):
        still.append(0.0, *[_frame(g, 0.0)] * 3)
        still.append(1.0, *[_frame(g, 0.0)] * 3)


@pytest.mark.parametrize("compress", [False, True])
def test_the_store_is_zarr_v3_on_disk(tmp_path, compress):
    """Metadata documents and chunk files as the Zarr v3 format defines them: a group
    document, one document per array, a regular grid, ``c/t/y/x`` chunk keys, the
    ``bytes`` codec (little endian) and optionally ``gzip`` after it."""
    g = Grid(ny=5, nx=7, dx=2.0)
    rng = np.random.default_rng(0)
    frames = [rng.uniform(0, 1, g.shape).astype(np.float32) for _ in range(3)]
    w = ZarrWriter(tmp_path / "r.zarr", g, 4, {"scheme": "local_inertial", "dx": 2.0},
                   compress=compress)
    w.write_bed(frames[0] + 10)
    w.write_static("channel_width", frames[1])
    for i, f in enumerate(frames):
        w.append(60.0 * i, f, -f, 2 * f)
    w.finalize({"mass_max_rel_error": 1e-9})
    root = tmp_path / "r.zarr"
    grp = json.loads((root / "zarr.json").read_text())
    assert grp == {"zarr_format": 3, "node_type": "group",
                   "attributes": {"scheme": "local_inertial", "dx": 2.0,
                                  "mass_max_rel_error": 1e-9, "n_frames": 3,
                                  "n_frames_allocated": 4}}
    meta = json.loads((root / "depth" / "zarr.json").read_text())
    assert meta["zarr_format"] == 3 and meta["node_type"] == "array"
    assert meta["shape"] == [4, 5, 7] and meta["data_type"] == "float32"
    assert meta["chunk_grid"] == {"name": "regular", "configuration": {"chunk_shape": [1, 5, 7]}}
    assert meta["chunk_key_encoding"] == {"name": "default", "configuration": {"separator": "/"}}
    assert meta["codecs"][0] == {"name": "bytes", "configuration": {"endian": "little"}}
    assert meta["dimension_names"] == ["time", "y", "x"]
    for i, f in enumerate(frames):
        raw = (root / "depth" / "c" / str(i) / "0" / "0").read_bytes()
        if compress:
            raw = gzip.decompress(raw)
        assert raw == f.astype("<f4").tobytes()
    assert not (root / "depth" / "c" / "3").exists()  # never written: reads as fill
    st = open_store(root)
    assert np.array_equal(st["v"].frame(2), 2 * frames[2])
    assert np.array_equal(st["depth"].frame(3), np.zeros(g.shape, np.float32))
    assert list(st["time"].read()) == [0.0, 60.0, 120.0, 0.0]
    assert np.array_equal(st["bed"].read(), frames[0] + 10)
    assert st["channel_width"].dims == ("y", "x")


# --- frames from oracle runs (the reference's export fixtures) ------------------------
def velocities(h, qx, qy, hu=None, hv=None):
    """state.py:232-256 restated: u, v from the faces (or momentum), 0 where dry."""
    wet = h >= H_DRY
    u = np.zeros_like(h)
    v = np.zeros_like(h)
    if hu is not None:
        np.divide(hu, h, out=u, where=wet)
        np.divide(hv, h, out=v, where=wet)
    else:
        np.divide(0.5 * (qx[:, :-1] + qx[:, 1:]), h, out=u, where=wet)
        np.divide(0.5 * (qy[:-1, :] + qy[1:, :]), h, out=v, where=wet)
    return u.astype(np.float32), v.astype(np.float32)


def _oracle_run(orc, path, bed, *, dx, end, every, rain_mm_hr, rain_s, dt_max, channels=None,
                sediment=None):
    """run.py:347-400 over the oracle: baseline frame, then one per output tick."""
    st = orc.State(bed, dx, depth=0.0, manning=0.035)
    if channels is not None:
        st.arm_channels(*channels)
    g = Grid(ny=bed.shape[0], nx=bed.shape[1], dx=dx)
    n_frames = int(round(end / every)) + 1
    w = ZarrWriter(path, g, n_frames, {"scheme": "local_inertial", "dx": dx, "crs": ""},
                   bed_change=sediment is not None)
    w.write_bed(bed)
    if channels is not None:
        w.write_static("channel_width", st.ch_w)
        w.write_static("channel_depth", st.ch_d)
    bc = (lambda: sediment) if sediment is not None else (lambda: None)
    w.append(0.0, st.h, *velocities(st.h, st.qx, st.qy), bc())
    rain = rain_mm_hr / 1000.0 / 3600.0
    sched = MultiRateScheduler(end_time=end, output_every=every, events=[rain_s])
    for tick in sched.ticks(lambda: orc.li_compute_dt(st, 0.7, dt_max)):
        orc.li_step(st, tick.dt, rain=rain if tick.t0 < rain_s - 1e-6 else 0.0)
        if tick.is_output:
            w.append(tick.t1, st.h, *velocities(st.h, st.qx, st.qy), bc())
    w.finalize()
    return path


def _bowl():
    ny, nx = 16, 20  # non-square: catches a transpose
    yy, xx = np.mgrid[0:ny, 0:nx]
    return (((yy - ny / 2) ** 2 + (xx - nx / 2) ** 2) * 0.02).astype(np.float32)


@pytest.fixture()
def bowl_store(orc, tmp_path):
    return _oracle_run(orc, tmp_path / "r.zarr", _bowl(), dx=20.0, end=600.0, every=200.0,
                       rain_mm_hr=100.0, rain_s=300.0, dt_max=10.0)


def _channel_store(orc, tmp_path):
    ny, nx = 16, 20
    yy, xx = np.mgrid[0:ny, 0:nx]
    bed = (0.4 * yy + 0.05 * (xx - nx / 2) ** 2).astype(np.float32)
    w = np.zeros((ny, nx), np.float32)
    d = np.zeros((ny, nx), np.float32)
    w[:, nx // 2] = 8.0
    d[:, nx // 2] = 1.0
    return _oracle_run(orc, tmp_path / "chan.zarr", bed, dx=20.0, end=600.0, every=150.0,
                       rain_mm_hr=150.0, rain_s=600.0, dt_max=5.0, channels=(w, d))


def test_frames_roundtrip_store_bytes(bowl_store, tmp_path):
    out = tmp_path / "frames"
    export_frames(bowl_store, out)
    st = open_store(bowl_store)
    ny, nx = st["depth"].shape[1:]
    m = json.loads((out / "manifest.json").read_text())
    assert m["n_frames"] == 4 and m["fields"] == ["depth"]
    assert m["grid"] == {"width": nx, "height": ny}
    assert m["tile_grid"]["cols"] == 1 and m["tile_grid"]["rows"] == 1
    assert "tiles" not in m["frames"][0]
    assert [f["time"] for f in m["frames"]] == [0.0, 200.0, 400.0, 600.0]
    for fr in m["frames"]:
        raw = np.fromfile(out / fr["files"]["depth"], dtype="<f4").reshape(ny, nx)
        expected = st["depth"].frame(fr["index"])
        assert np.array_equal(raw, expected)
        assert fr["depth"]["min"] == float(expected.min())
        assert fr["depth"]["max"] == float(expected.max())
    assert float(st["depth"].frame(3).max()) > H_DRY  # the storm wetted the bowl


def test_global_stats_are_robust_and_bounded(bowl_store, tmp_path):
    out = tmp_path / "frames"
    export_frames(bowl_store, out)
    g = json.loads((out / "manifest.json").read_text())["global"]["depth"]
    depth = open_store(bowl_store)["depth"].read()
    assert g["max"] == float(depth.max())
    assert 0.0 <= g["p50"] <= g["p99"] <= g["max"] and g["min"] >= 0.0
    wet = depth[depth >= H_DRY].astype(np.float64)
    assert g["p99"] == pytest.approx(float(np.percentile(wet, 99.0)), rel=1e-12)


def test_export_via_cli_entrypoint(bowl_store, tmp_path):
    viewer_export.main([str(bowl_store), str(tmp_path / "cli")])
    assert (tmp_path / "cli" / "manifest.json").exists()


def _reassemble(out, tiles, names, shape):
    rebuilt = np.zeros(shape, np.float32)
    for t, name in zip(tiles, names):
        block = np.fromfile(out / name, dtype="<f4").reshape(t["height"], t["width"])
        rebuilt[t["y"]:t["y"] + t["height"], t["x"]:t["x"] + t["width"]] = block
    return rebuilt


def test_large_frames_split_into_a_tile_grid_that_reassembles(bowl_store, tmp_path):
    out = tmp_path / "tiled"
    export_frames(bowl_store, out, tile_size=8)
    st = open_store(bowl_store)
    ny, nx = st["depth"].shape[1:]
    m = json.loads((out / "manifest.json").read_text())
    grid = m["tile_grid"]
    assert (grid["rows"], grid["cols"], grid["size"]) == (2, 3, 8)
    assert [t["width"] for t in grid["tiles"][:3]] == [8, 8, 4]
    assert sum(t["width"] * t["height"] for t in grid["tiles"]) == ny * nx
    for fr in m["frames"]:
        assert fr["files"] == {}
        assert np.array_equal(_reassemble(out, grid["tiles"], fr["tiles"]["depth"], (ny, nx)),
                              st["depth"].frame(fr["index"]))
    static = m["static"]
    assert static["files"] == {}
    assert np.array_equal(_reassemble(out, grid["tiles"], static["tiles"]["bed"], (ny, nx)),
                          st["bed"].read())


def test_static_bed_ships_with_the_frames(bowl_store, tmp_path):
    out = tmp_path / "frames"
    export_frames(bowl_store, out)
    m = json.loads((out / "manifest.json").read_text())
    bed = open_store(bowl_store)["bed"].read()
    static = m["static"]
    assert static["fields"] == ["bed"] and "channel" not in static
    assert np.array_equal(np.fromfile(out / static["files"]["bed"], dtype="<f4").reshape(bed.shape),
                          bed)
    assert static["bed"] == {"min": float(bed.min()), "max": float(bed.max())}
    assert "morphology" not in m
    assert not any("channel" in p.name for p in out.iterdir())


def test_manifest_carries_the_mosaic_assembly_record(bowl_store, tmp_path):
    grp = json.loads((bowl_store / "zarr.json").read_text())
    grp["attributes"].update({"coarsen": 2, "domain": {"gap_cells": 7, "fill_value": 12.5}})
    (bowl_store / "zarr.json").write_text(json.dumps(grp))
    m = json.loads(export_frames(bowl_store, tmp_path / "f").read_text())
    assert m["coarsen"] == 2 and m["domain"] == {"gap_cells": 7, "fill_value": 12.5}


def test_a_morphological_store_declares_its_terrain_is_the_starting_bed(orc, tmp_path):
    bed = _bowl()
    dz = np.linspace(-0.3, 0.1, bed.size, dtype=np.float32).reshape(bed.shape)
    path = _oracle_run(orc, tmp_path / "sed.zarr", bed, dx=20.0, end=600.0, every=300.0,
                       rain_mm_hr=120.0, rain_s=600.0, dt_max=5.0, sediment=dz)
    out = tmp_path / "frames"
    m = json.loads(export_frames(path, out).read_text())
    morph = m["morphology"]
    assert morph["static_bed"] == "initial"
    assert morph["bed_change"]["min"] == float(dz.min()) < 0.0
    assert morph["bed_change"]["max"] == float(dz.max())
    assert morph["bed_change"]["time"] == 600.0
    assert not any("bed_change" in p.name for p in out.iterdir())


def test_the_drawn_surface_is_the_storage_curve_overbank_and_the_bank_below_it():
    """test_viewer_export.py:296-324."""
    dx = 20.0
    z, w, d = np.full(9, 100.0), np.full(9, 8.0), np.full(9, 1.0)
    h_bf = 0.4
    h = np.array([0.0, 0.05, 0.2, 0.399, h_bf, 0.401, 0.6, 1.0, 3.0])
    drawn, true = render_eta(h, z, w, d, dx), eta_from_h(h, z, w, d, dx)
    over = h > h_bf
    assert np.array_equal(drawn[over], true[over])
    assert np.all(drawn[~over] == z[~over])
    assert np.all(drawn >= true - 1e-12)
    assert drawn[3] == pytest.approx(drawn[5], abs=2e-3)
    assert drawn[1] - true[1] == pytest.approx(d[1] - h[1] * dx / w[1], rel=1e-12)
    zeros = np.zeros(3)
    zz, hh = np.array([1.0, 250.0, -3.5]), np.array([0.0, 0.001, 2.25])
    assert np.array_equal(render_eta(hh, zz, zeros, zeros, 20.0), zz + hh)
    assert np.array_equal(render_eta(hh, zz, np.full(3, 8.0), zeros, 20.0), zz + hh)


def test_channel_geometry_ships_and_the_offset_is_measured(orc, tmp_path):
    """test_viewer_export.py:343-432: the geometry rides the static entry (tiled too),
    and the manifest's in-bank offset is the worst over all frames."""
    path = _channel_store(orc, tmp_path)
    out = tmp_path / "frames"
    m = json.loads(export_frames(path, out).read_text())
    st = open_store(path)
    static = m["static"]
    assert static["fields"] == ["bed", "channel_width", "channel_depth"]
    for name in ("channel_width", "channel_depth"):
        got = np.fromfile(out / static["files"][name], dtype="<f4")
        assert np.array_equal(got.reshape(16, 20), st[name].read())
    chan = static["channel"]
    w, d, z = st["channel_width"].read(), st["channel_depth"].read(), st["bed"].read()
    assert chan["cells"] == 16 and chan["width_max_m"] == 8.0 and chan["depth_max_m"] == 1.0
    has = (w > 0) & (d > 0)
    h_bf = np.where(has, w * d / 20.0, 0.0)
    best = (0.0, 0, -1)
    for i in range(m["n_frames"]):
        h = st["depth"].frame(i)
        in_bank = has & (h >= H_DRY) & (h <= h_bf)
        if in_bank.any():
            gap = float((render_eta(h, z, w, d, 20.0) - eta_from_h(h, z, w, d, 20.0))[in_bank].max())
            if gap > best[0]:
                best = (gap, int(in_bank.sum()), i)
    assert best[1] > 0
    assert chan["in_bank_offset_m"] == pytest.approx(best[0], rel=1e-12)
    assert (chan["in_bank_cells"], chan["frame"]) == best[1:]
    assert 0.0 < chan["in_bank_offset_m"] <= chan["depth_max_m"]
    tiled = tmp_path / "tiled"
    mt = json.loads(export_frames(path, tiled, tile_size=8).read_text())
    for name in ("bed", "channel_width", "channel_depth"):
        assert np.array_equal(_reassemble(tiled, mt["tile_grid"]["tiles"],
                                          mt["static"]["tiles"][name], (16, 20)), st[name].read())


def test_untiled_export_bytes_do_not_depend_on_the_tile_size(bowl_store, tmp_path):
    a, b = tmp_path / "a", tmp_path / "b"
    export_frames(bowl_store, a)
    export_frames(bowl_store, b, tile_size=4096)
    for fr in json.loads((a / "manifest.json").read_text())["frames"]:
        name = fr["files"]["depth"]
        assert (a / name).read_bytes() == (b / name).read_bytes()


def test_an_early_stop_exports_only_the_written_frames(tmp_path):
    g = _grid()
    w = ZarrWriter(tmp_path / "r.zarr", g, 5, {"dx": g.dx})
    w.write_bed(_frame(g, 1.0))
    w.append(0.0, _frame(g, 0.2), _frame(g, 0.0), _frame(g, 0.0))
    w.append(10.0, _frame(g, 0.3), _frame(g, 0.0), _frame(g, 0.0))
    w.finalize()
    m = json.loads(export_frames(tmp_path / "r.zarr", tmp_path / "f").read_text())
    assert m["n_frames"] == 2 and len(m["frames"]) == 2
