"""Per-frame viewer tiles from the result store (reference solver/io/viewer_export.py).

The viewer cannot read Zarr, so a finished store is exported as one raw
little-endian float32 file per frame and field (row-major, the terrain loader's
orientation), split into a row-major tile grid when a frame exceeds ``tile_size``,
plus ``manifest.json``: the tile geometry once, each frame's files and min/max, and
robust colour ranges over the wet cells of all frames (viewer_export.py:316-452).
The run's bed, and its sub-grid channel geometry when it had one, ship through the
same tile layout as static fields; a channelled run's manifest also declares how far
the drawn river surface sits above the true one (``render_eta``), and a
morphological run's declares that the shipped terrain is the initial bed.

The manifest's keys, the file names and the byte layout are the reference's, so its
viewer reads these exports unchanged.  The store is read with
:func:`rbs.io.zarr_store.open_store`.
"""

from __future__ import annotations

import argparse
import json
from pathlib import Path

import numpy as np

from rbs.core.channels import eta_from_h
from rbs.core.grid import H_DRY
from rbs.io.zarr_store import open_store

DEFAULT_TILE_SIZE = 512  # viewer_export.py:81-84
_PCTL_SAMPLE_CAP = 4_000_000  # wet samples kept per frame for the global percentiles


def render_eta(h, z, w, d, dx: float) -> np.ndarray:
    """The surface the viewer draws (viewer_export.py:94-120): the storage curve above
    bank full, the bank itself below it (the true surface is under a floodplain bed
    the rendered terrain has no trench in), ``z + h`` without a channel."""
    h, z, w, d = (np.asarray(v, dtype=np.float64) for v in (h, z, w, d))
    has = (w > 0.0) & (d > 0.0)
    h_bf = np.where(has, w * d / dx, 0.0)
    return np.where(has, z + np.maximum(h - h_bf, 0.0), z + h)


def _robust_stats(wet: np.ndarray) -> dict:
    if wet.size == 0:
        return {"min": 0.0, "max": 0.0, "p50": 0.0, "p99": 0.0}
    return {"min": float(wet.min()), "max": float(wet.max()),
            "p50": float(np.percentile(wet, 50.0)), "p99": float(np.percentile(wet, 99.0))}


def tile_layout(ny: int, nx: int, tile_size: int) -> list[dict]:
    """Row-major tiles covering (ny, nx); edge tiles clipped, not padded."""
    out = []
    for r, y0 in enumerate(range(0, ny, tile_size)):
        for c, x0 in enumerate(range(0, nx, tile_size)):
            out.append({"row": r, "col": c, "x": x0, "y": y0,
                        "width": min(tile_size, nx - x0), "height": min(tile_size, ny - y0)})
    return out


def _write_field(arr, out_dir: Path, layout, *, field: str, stem: str, tiled: bool) -> dict:
    if not tiled:
        name = f"{stem}.raw"
        (out_dir / name).write_bytes(np.ascontiguousarray(arr, dtype="<f4").tobytes())
        return {"files": {field: name}}
    names = []
    for t in layout:
        block = np.ascontiguousarray(arr[t["y"]:t["y"] + t["height"], t["x"]:t["x"] + t["width"]],
                                     dtype="<f4")
        name = f"{stem}_r{t['row']:02d}_c{t['col']:02d}.raw"
        (out_dir / name).write_bytes(block.tobytes())
        names.append(name)
    return {"files": {}, "tiles": {field: names}}


def _static_entry(store, out_dir, layout, tiled, channels):
    if "bed" not in store:
        return None
    bed = np.ascontiguousarray(store["bed"].read(), dtype="<f4")
    entry = _write_field(bed, out_dir, layout, field="bed", stem="bed", tiled=tiled)
    fields = ["bed"]
    if channels is not None:
        for name, arr in zip(("channel_width", "channel_depth"), channels):
            payload = _write_field(arr, out_dir, layout, field=name, stem=name, tiled=tiled)
            entry["files"].update(payload["files"])
            if "tiles" in payload:
                entry.setdefault("tiles", {}).update(payload["tiles"])
            fields.append(name)
    entry["fields"] = fields
    entry["bed"] = {"min": float(bed.min()), "max": float(bed.max())}
    return entry


def export_frames(store_path, out_dir, *, field: str = "depth", h_dry: float = H_DRY,
                  tile_size: int = DEFAULT_TILE_SIZE) -> Path:
    """Export ``field`` of every written frame as viewer tiles; returns the manifest path."""
    store = open_store(store_path)
    out_dir = Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    arr3 = store[field]
    n_alloc = arr3.shape[0]
    n_frames = min(int(store.attrs.get("n_frames", n_alloc)), n_alloc)
    ny, nx = arr3.shape[1], arr3.shape[2]
    times = [float(t) for t in store["time"].read()]
    layout = tile_layout(ny, nx, max(int(tile_size), 1))
    tiled = len(layout) > 1
    rows = max(t["row"] for t in layout) + 1
    cols = max(t["col"] for t in layout) + 1
    dx = float(store.attrs.get("dx", 1.0))

    channels = None
    if "channel_width" in store and "channel_depth" in store:
        channels = (np.ascontiguousarray(store["channel_width"].read(), dtype="<f4"),
                    np.ascontiguousarray(store["channel_depth"].read(), dtype="<f4"))
    measure = channels is not None and field == "depth"
    offset = {"offset": 0.0, "cells": 0, "frame": -1}
    if measure:
        cw, cd = channels
        has = (cw > 0) & (cd > 0)
        h_bf = np.where(has, cw * cd / dx, 0.0)
        bed64 = np.asarray(store["bed"].read(), dtype=np.float64)

    frames, wet_samples, global_max = [], [], 0.0
    for i in range(n_frames):
        arr = np.ascontiguousarray(arr3.frame(i), dtype="<f4")
        payload = _write_field(arr, out_dir, layout, field=field, stem=f"f{i:04d}_{field}",
                               tiled=tiled)
        if measure:
            in_bank = has & (arr >= h_dry) & (arr <= h_bf)
            if in_bank.any():
                gap = (render_eta(arr, bed64, cw, cd, dx) - eta_from_h(arr, bed64, cw, cd, dx))
                worst = float(gap[in_bank].max())
                if worst > offset["offset"]:
                    offset = {"offset": worst, "cells": int(in_bank.sum()), "frame": i}
        fmin, fmax = float(arr.min()), float(arr.max())
        global_max = max(global_max, fmax)
        wet = arr[arr >= h_dry].astype(np.float64, copy=False)
        if wet.size > _PCTL_SAMPLE_CAP:
            wet = wet[::int(np.ceil(wet.size / _PCTL_SAMPLE_CAP))]
        wet_samples.append(wet)
        frames.append({"index": i, "time": times[i], **payload, field: {"min": fmin, "max": fmax}})

    stats = _robust_stats(np.concatenate(wet_samples) if wet_samples else np.empty(0))
    stats["max"] = global_max
    manifest = {
        "dx": dx,
        "crs": str(store.attrs.get("crs", "")),
        "scheme": str(store.attrs.get("scheme", "")),
        "coarsen": int(store.attrs.get("coarsen", 1)),
        "grid": {"width": nx, "height": ny},
        "tile_grid": {"cols": cols, "rows": rows, "size": int(tile_size), "tiles": layout},
        "fields": [field],
        "h_dry": float(h_dry),
        "n_frames": n_frames,
        "global": {field: stats},
        "frames": frames,
    }
    static = _static_entry(store, out_dir, layout, tiled, channels)
    if static is not None:
        if measure:
            cw, cd = channels
            chan_has = (cw > 0) & (cd > 0)
            static["channel"] = {
                "cells": int(chan_has.sum()),
                "width_max_m": float(cw.max()),
                "depth_max_m": float(cd.max()),
                "dx_m": dx,
                "in_bank_offset_m": offset["offset"],
                "in_bank_cells": offset["cells"],
                "frame": offset["frame"],
                "note": ("below bank full the river is drawn at the bank; its true surface "
                         "lies under the floodplain bed, which the terrain carries no trench for"),
            }
        manifest["static"] = static
    if "bed_change" in store and n_frames > 0:
        dz = np.asarray(store["bed_change"].frame(n_frames - 1), dtype=np.float64)
        manifest["morphology"] = {
            "bed_change": {"min": float(dz.min()), "max": float(dz.max()),
                           "frame": n_frames - 1, "time": times[n_frames - 1]},
            "static_bed": "initial",
            "note": ("the terrain is the bed at t=0; the run moved it (bed_change in the "
                     "result store) and the viewer does not animate terrain"),
        }
    domain = store.attrs.get("domain")
    if isinstance(domain, dict):
        manifest["domain"] = dict(domain)
    path = out_dir / "manifest.json"
    path.write_text(json.dumps(manifest, indent=2), encoding="utf-8")
    return path


def main(argv=None) -> None:
    p = argparse.ArgumentParser(description="Export per-frame viewer tiles from a result store.")
    p.add_argument("store", help="result store (Zarr v3 directory)")
    p.add_argument("out_dir", help="output frames/ directory")
    p.add_argument("--field", default="depth")
    args = p.parse_args(argv)
    print(f"viewer export: {export_frames(args.store, args.out_dir, field=args.field)}")


if __name__ == "__main__":
    main()
