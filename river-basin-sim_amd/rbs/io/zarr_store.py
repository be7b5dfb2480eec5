"""The canonical result store: a Zarr v3 group written frame by frame (reference
solver/io/zarr_writer.py:33-154).

Layout (zarr_writer.py:8-21)::

    results.zarr/
      zarr.json  group attributes: crs, dx, units, scheme, ..., mass series
      time       (T,)      float64
      depth      (T, Y, X) float32, one frame per chunk
      u, v       (T, Y, X) float32
      bed        (Y, X)    float32, the initial bed
      bed_change (T, Y, X) float32, only when morphology ran
      <static>   (Y, X)    float32, e.g. channel_width / channel_depth

Every array carries ``dimension_names``.  The reference writes through zarr-python;
this writer emits the Zarr v3 format itself (metadata documents and chunk files, a
regular chunk grid, the default ``c/i/j/k`` chunk keys, the ``bytes`` codec with an
optional ``gzip`` after it), so the store needs no zarr installation to write and
opens with any Zarr v3 reader.  :func:`open_store` is a minimal reader of what this
writer produces.
"""

from __future__ import annotations

import gzip
import json
import os
from pathlib import Path

import numpy as np

_DTYPES = {"f4": ("float32", np.dtype("<f4")), "f8": ("float64", np.dtype("<f8"))}


def _codecs(compress: bool) -> list[dict]:
    codecs = [{"name": "bytes", "configuration": {"endian": "little"}}]
    if compress:
        codecs.append({"name": "gzip", "configuration": {"level": 1}})
    return codecs


def _write_json(path: Path, doc: dict) -> None:
    tmp = path.with_suffix(".json.tmp")
    tmp.write_text(json.dumps(doc, indent=2, default=_json_default), encoding="utf-8")
    os.replace(tmp, path)


def _json_default(o):
    if isinstance(o, (np.floating, np.integer)):
        return o.item()
    if isinstance(o, np.ndarray):
        return o.tolist()
    raise TypeError(f"not JSON serialisable: {type(o).__name__}")


class _Array:
    """One Zarr v3 array of a regular chunk grid, written chunk by chunk."""

    def __init__(self, root: Path, name: str, shape, chunks, dtype: str, dims, compress: bool):
        self.path = root / name
        self.path.mkdir(parents=True, exist_ok=True)
        self.shape = tuple(int(s) for s in shape)
        self.chunks = tuple(int(c) for c in chunks)
        self.dtype = _DTYPES[dtype][1]
        self.compress = compress
        _write_json(self.path / "zarr.json", {
            "zarr_format": 3,
            "node_type": "array",
            "shape": list(self.shape),
            "data_type": _DTYPES[dtype][0],
            "chunk_grid": {"name": "regular", "configuration": {"chunk_shape": list(self.chunks)}},
            "chunk_key_encoding": {"name": "default", "configuration": {"separator": "/"}},
            "fill_value": 0.0,
            "codecs": _codecs(compress),
            "attributes": {},
            "dimension_names": list(dims),
            "storage_transformers": [],
        })

    def write_chunk(self, index: tuple[int, ...], block: np.ndarray) -> None:
        """Write one whole chunk (edge chunks are padded to the chunk shape).  A block
        that already is a contiguous little-endian chunk is written from its own
        buffer, without a copy."""
        inner = tuple(min(c, s - i * c) for c, s, i in zip(self.chunks, self.shape, index))
        block = np.asarray(block)
        if inner == self.chunks and block.dtype == self.dtype and block.flags["C_CONTIGUOUS"]:
            data = memoryview(block.reshape(self.chunks)).cast("B")
        else:
            full = np.zeros(self.chunks, dtype=self.dtype)
            full[tuple(slice(0, n) for n in inner)] = np.asarray(block, dtype=self.dtype).reshape(inner)
            data = memoryview(full).cast("B")
        if self.compress:
            data = gzip.compress(data, compresslevel=1, mtime=0)
        key = self.path.joinpath("c", *[str(i) for i in index])
        key.parent.mkdir(parents=True, exist_ok=True)
        with open(key, "wb") as f:
            f.write(data)

    def write_all(self, values: np.ndarray) -> None:
        values = np.asarray(values, dtype=self.dtype)
        if values.shape != self.shape:
            raise ValueError(f"shape {values.shape} != {self.shape}")
        grid = [range(-(-s // c)) for s, c in zip(self.shape, self.chunks)]
        for idx in np.ndindex(*[len(g) for g in grid]):
            sl = tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, self.chunks, self.shape))
            self.write_chunk(idx, values[sl])


class ZarrWriter:
    """Incremental writer of the canonical store (zarr_writer.py:33-154): ``n_frames``
    preallocated along time, frames appended by index, attributes merged at
    ``finalize``; ``bed_change`` exists exactly when the run has morphology."""

    def __init__(self, path, grid, n_frames: int, attrs: dict, *, bed_change: bool = False,
                 compress: bool = False, parallel: bool = True):
        self.path = Path(path)
        self._pool = None
        if parallel:
            from concurrent.futures import ThreadPoolExecutor

            self._pool = ThreadPoolExecutor(max_workers=4, thread_name_prefix="rbs-store")
        self.grid = grid
        self.n_frames = int(n_frames)
        self._i = 0
        self.compress = compress
        self.path.mkdir(parents=True, exist_ok=True)
        self.attrs = dict(attrs)
        self._write_group()
        ny, nx = grid.shape
        tyx = ("time", "y", "x")
        frame = (1, ny, nx)  # one timestep per chunk (zarr_writer.py:55)
        mk = lambda name, shape, chunks, dt, dims: _Array(self.path, name, shape, chunks, dt,  # noqa: E731
                                                          dims, compress)
        self._depth = mk("depth", (self.n_frames, ny, nx), frame, "f4", tyx)
        self._u = mk("u", (self.n_frames, ny, nx), frame, "f4", tyx)
        self._v = mk("v", (self.n_frames, ny, nx), frame, "f4", tyx)
        self._time = mk("time", (self.n_frames,), (self.n_frames,), "f8", ("time",))
        self._times = np.zeros(self.n_frames, np.float64)
        self._bed = mk("bed", (ny, nx), (ny, nx), "f4", ("y", "x"))
        self._bed_change = mk("bed_change", (self.n_frames, ny, nx), frame, "f4", tyx) \
            if bed_change else None

    def _write_group(self) -> None:
        _write_json(self.path / "zarr.json",
                    {"zarr_format": 3, "node_type": "group", "attributes": self.attrs})

    def write_bed(self, bed: np.ndarray) -> None:
        """The static bed, once (zarr_writer.py:93-95)."""
        self._bed.write_all(np.ascontiguousarray(bed, dtype=np.float32))

    def write_static(self, name: str, field: np.ndarray) -> None:
        """An extra static (Y, X) field, e.g. channel geometry (zarr_writer.py:97-108)."""
        ny, nx = self.grid.shape
        arr = _Array(self.path, name, (ny, nx), (ny, nx), "f4", ("y", "x"), self.compress)
        arr.write_all(np.ascontiguousarray(field, dtype=np.float32))

    def append(self, time: float, depth, u, v, bed_change=None) -> None:
        """Append one frame at the next index (zarr_writer.py:110-148); bed_change is
        required exactly when the store carries it."""
        if self._i >= self.n_frames:
            raise IndexError(f"more frames than preallocated ({self.n_frames})")
        if (bed_change is None) != (self._bed_change is None):
            raise ValueError(
                "bed_change must be passed exactly when the store carries it "
                f"(store: {'yes' if self._bed_change is not None else 'no'}, "
                f"frame: {'yes' if bed_change is not None else 'no'}); an omitted "
                "frame would store zeros, which reads as 'the bed did not move'")
        i = self._i
        jobs = [(self._depth, depth), (self._u, u), (self._v, v)]
        if self._bed_change is not None:
            jobs.append((self._bed_change, np.asarray(bed_change, dtype=np.float32)))
        if self._pool is not None:  # file writes release the GIL: the fields go side by side
            for fut in [self._pool.submit(a.write_chunk, (i, 0, 0), x) for a, x in jobs]:
                fut.result()
        else:
            for a, x in jobs:
                a.write_chunk((i, 0, 0), x)
        self._times[i] = float(time)
        self._time.write_chunk((0,), self._times)
        self._i += 1

    @property
    def frames_written(self) -> int:
        return self._i

    def finalize(self, extra_attrs: dict | None = None) -> None:
        """Merge final attributes and record the frame count (zarr_writer.py:150-154)."""
        if extra_attrs:
            self.attrs.update(extra_attrs)
        self.attrs["n_frames"] = self._i
        if self._i < self.n_frames:
            self.attrs["n_frames_allocated"] = self.n_frames
        self._write_group()
        if self._pool is not None:
            self._pool.shutdown(wait=True)
            self._pool = None


# --- reading --------------------------------------------------------------------------
class StoreArray:
    """A Zarr v3 array of a regular grid with the ``bytes`` (+ ``gzip``) codecs."""

    def __init__(self, path: Path):
        meta = json.loads((path / "zarr.json").read_text(encoding="utf-8"))
        if meta.get("zarr_format") != 3 or meta.get("node_type") != "array":
            raise ValueError(f"{path} is not a Zarr v3 array")
        self.path = path
        self.meta = meta
        self.shape = tuple(meta["shape"])
        self.chunks = tuple(meta["chunk_grid"]["configuration"]["chunk_shape"])
        self.dtype = np.dtype({"float32": "<f4", "float64": "<f8"}[meta["data_type"]])
        self.dims = tuple(meta.get("dimension_names") or ())
        names = [c["name"] for c in meta["codecs"]]
        if names[0] != "bytes" or any(n not in ("bytes", "gzip") for n in names):
            raise ValueError(f"unsupported codecs {names}")
        self._gzip = "gzip" in names
        self.fill = meta.get("fill_value", 0.0)

    def chunk(self, index: tuple[int, ...]) -> np.ndarray:
        key = self.path.joinpath("c", *[str(i) for i in index])
        if not key.exists():
            return np.full(self.chunks, self.fill, self.dtype)
        data = key.read_bytes()
        if self._gzip:
            data = gzip.decompress(data)
        return np.frombuffer(data, dtype=self.dtype).reshape(self.chunks)

    def read(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        counts = [-(-s // c) for s, c in zip(self.shape, self.chunks)]
        for idx in np.ndindex(*counts):
            sl = tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, self.chunks, self.shape))
            out[sl] = self.chunk(idx)[tuple(slice(0, x.stop - x.start) for x in sl)]
        return out

    def frame(self, i: int) -> np.ndarray:
        """Time index ``i`` of a (T, Y, X) array chunked one frame per chunk."""
        if len(self.shape) != 3 or self.chunks[0] != 1 or self.chunks[1:] != self.shape[1:]:
            return self.read()[i]
        return self.chunk((i, 0, 0))[0]


class Store:
    """A group written by :class:`ZarrWriter`: ``attrs`` and arrays by name."""

    def __init__(self, path):
        self.path = Path(path)
        meta = json.loads((self.path / "zarr.json").read_text(encoding="utf-8"))
        if meta.get("zarr_format") != 3 or meta.get("node_type") != "group":
            raise ValueError(f"{path} is not a Zarr v3 group")
        self.attrs = meta.get("attributes", {})
        self.arrays = {p.name: StoreArray(p) for p in sorted(self.path.iterdir())
                       if (p / "zarr.json").exists()}

    def __contains__(self, name: str) -> bool:
        return name in self.arrays

    def __getitem__(self, name: str) -> StoreArray:
        return self.arrays[name]


def open_store(path) -> Store:
    return Store(path)
