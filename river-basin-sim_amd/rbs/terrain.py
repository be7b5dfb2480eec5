"""Seeded synthetic inputs for the benchmark workloads.

The reference has no terrain generator (its demo beds are analytic); these build
the scenarios the benchmark names, deterministically from a seed:

* ``fbm_dem`` -- spectral fractional-Brownian surface, PSD ~ k^-(2H+2), rescaled
  to a given relief;
* ``carve_channels`` -- priority-flood pit filling, D8 flow accumulation, and
  channels cut 1..3 m deep (scaled by log drained area) where the drained area
  exceeds a threshold;
* ``voronoi_field`` -- piecewise-constant patches (e.g. Manning n);
* ``lognormal_field`` -- ``median * exp(sigma * G)`` with G a unit-variance
  Gaussian field smoothed to a correlation length;
* ``uniform_field`` -- i.i.d. uniform cells;
* ``reach_basin`` -- a coarse reach-scale basin whose rivers are carried as
  sub-grid channels sized from drained area by the reference pipeline's hydraulic
  geometry law (pipeline/channels.py:98-103, :767-820).
"""

from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

_LIB = Path(__file__).resolve().parent / "libterrain.so"
_lib = None


def _terrain_lib():
    global _lib
    if _lib is None:
        if not _LIB.exists():
            raise RuntimeError(f"{_LIB} is not built; run river-basin-sim_amd/build.py")
        L = ctypes.CDLL(str(_LIB))
        fp = ctypes.POINTER(ctypes.c_float)
        L.terrain_fill.argtypes = [ctypes.c_int32, ctypes.c_int32, fp, ctypes.c_float]
        L.terrain_d8_accumulate.argtypes = [ctypes.c_int32, ctypes.c_int32, fp, ctypes.c_float, fp]
        _lib = L
    return _lib


def fbm_dem(ny: int, nx: int, *, hurst: float = 0.8, relief: float = 200.0,
            seed: int = 0) -> np.ndarray:
    """Spectral fBm surface in [0, relief] metres."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((ny, nx))
    spec = np.fft.rfft2(noise)
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    amp = k ** (-(hurst + 1.0))
    amp[0, 0] = 0.0
    surf = np.fft.irfft2(spec * amp, s=(ny, nx))
    surf -= surf.min()
    surf *= relief / max(surf.max(), 1e-12)
    return surf.astype(np.float32)


def d8_area(dem: np.ndarray, dx: float) -> tuple[np.ndarray, np.ndarray]:
    """Pit-filled surface and D8 drained area (cells, including the cell itself)."""
    L = _terrain_lib()
    ny, nx = dem.shape
    z = np.ascontiguousarray(dem, dtype=np.float32).copy()
    fp = ctypes.POINTER(ctypes.c_float)
    if L.terrain_fill(ny, nx, z.ctypes.data_as(fp), ctypes.c_float(1e-4)) != 0:
        raise MemoryError("terrain_fill")
    acc = np.empty_like(z)
    if L.terrain_d8_accumulate(ny, nx, z.ctypes.data_as(fp), ctypes.c_float(dx),
                               acc.ctypes.data_as(fp)) != 0:
        raise MemoryError("terrain_d8_accumulate")
    return z, acc


def carve_channels(dem: np.ndarray, dx: float, *, area_threshold_m2: float = 1.0e6,
                   depth_min: float = 1.0, depth_max: float = 3.0) -> np.ndarray:
    """Fill pits, accumulate D8 flow, cut channels depth_min..depth_max deep."""
    L = _terrain_lib()
    ny, nx = dem.shape
    z = np.ascontiguousarray(dem, dtype=np.float32).copy()
    fp = ctypes.POINTER(ctypes.c_float)
    if L.terrain_fill(ny, nx, z.ctypes.data_as(fp), ctypes.c_float(1e-4)) != 0:
        raise MemoryError("terrain_fill")
    acc = np.empty_like(z)
    if L.terrain_d8_accumulate(ny, nx, z.ctypes.data_as(fp), ctypes.c_float(dx),
                               acc.ctypes.data_as(fp)) != 0:
        raise MemoryError("terrain_d8_accumulate")
    thr = max(area_threshold_m2 / (dx * dx), 2.0)
    chan = acc >= thr
    if chan.any():
        la = np.log(acc[chan])
        lo, hi = np.log(thr), max(float(la.max()), np.log(thr) + 1e-6)
        depth = depth_min + (depth_max - depth_min) * (la - lo) / (hi - lo)
        z[chan] -= depth.astype(np.float32)
    return z


def voronoi_field(ny: int, nx: int, *, sites: int = 512, lo: float = 0.02, hi: float = 0.10,
                  seed: int = 0) -> np.ndarray:
    """Nearest-site patches, each uniform in [lo, hi]."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    pts = np.column_stack([rng.uniform(0, ny, sites), rng.uniform(0, nx, sites)])
    vals = rng.uniform(lo, hi, sites).astype(np.float32)
    tree = cKDTree(pts)
    out = np.empty((ny, nx), np.float32)
    cols = np.arange(nx) + 0.5
    step = max(1, (1 << 22) // nx)
    for r0 in range(0, ny, step):
        r1 = min(ny, r0 + step)
        yy, xx = np.meshgrid(np.arange(r0, r1) + 0.5, cols, indexing="ij")
        _, idx = tree.query(np.column_stack([yy.ravel(), xx.ravel()]))
        out[r0:r1] = vals[idx].reshape(r1 - r0, nx)
    return out


def gaussian_field(ny: int, nx: int, corr_cells: float, *, seed: int = 0) -> np.ndarray:
    """Unit-variance Gaussian field smoothed with a Gaussian kernel of width corr_cells."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((ny, nx))
    ky = np.fft.fftfreq(ny)[:, None]
    kx = np.fft.rfftfreq(nx)[None, :]
    kern = np.exp(-2.0 * (np.pi * corr_cells) ** 2 * (kx * kx + ky * ky))
    g = np.fft.irfft2(np.fft.rfft2(noise) * kern, s=(ny, nx))
    g -= g.mean()
    g /= max(g.std(), 1e-12)
    return g


def lognormal_field(ny: int, nx: int, *, median: float, sigma: float, corr_cells: float,
                    seed: int = 0) -> np.ndarray:
    return (median * np.exp(sigma * gaussian_field(ny, nx, corr_cells, seed=seed))).astype(
        np.float32)


def uniform_field(ny: int, nx: int, lo: float, hi: float, *, seed: int = 0) -> np.ndarray:
    return np.random.default_rng(seed).uniform(lo, hi, (ny, nx)).astype(np.float32)


MM_H = 1.0 / 1000.0 / 3600.0  # mm/h -> m/s


def rain_on_terrain(n: int, *, dx: float = 10.0, seed: int = 20261203) -> dict:
    """The rain-on-terrain basin: n x n, fBm (H 0.8, 200 m relief) with D8 channels,
    Voronoi Manning in [0.02, 0.10], lognormal rain (median 15 mm/h, sigma 0.6,
    2 km correlation), infiltration U[0, 10] mm/h.  Rates are returned in m/s."""
    dem = fbm_dem(n, n, hurst=0.8, relief=200.0, seed=seed)
    bed = carve_channels(dem, dx)
    manning = voronoi_field(n, n, sites=512, lo=0.02, hi=0.10, seed=seed + 1)
    rain = lognormal_field(n, n, median=15.0, sigma=0.6, corr_cells=2000.0 / dx,
                           seed=seed + 2) * np.float32(MM_H)
    infil = uniform_field(n, n, 0.0, 10.0, seed=seed + 3) * np.float32(MM_H)
    return {"bed": bed, "manning": manning, "rain": rain.astype(np.float32),
            "infil": infil.astype(np.float32), "dx": dx}


def hydraulic_geometry(area_km2: np.ndarray, dx: float, *, width_coef: float = 8.0,
                       width_exp: float = 0.5, depth_coef: float = 0.27,
                       depth_exp: float = 0.3, min_area_km2: float = 1.0):
    """Channel width and bank-full depth from drained area, w = 8 A^0.5 m and
    d = 0.27 A^0.3 m for A >= 1 km^2 (the reference pipeline's defaults), widths
    capped at the largest float32 not above dx; "no channel" is 0 in both."""
    area = np.asarray(area_km2, dtype=np.float64)
    river = area >= min_area_km2
    width = np.where(river, width_coef * np.power(np.maximum(area, 0.0), width_exp), 0.0)
    depth = np.where(river, depth_coef * np.power(np.maximum(area, 0.0), depth_exp), 0.0)
    cap = np.float32(dx)
    if float(cap) > float(dx):
        cap = np.nextafter(cap, np.float32(0.0))
    width = np.minimum(width, float(cap))
    depth = np.where(width > 0.0, depth, 0.0)
    return width.astype(np.float32), depth.astype(np.float32)


def reach_basin(n: int, *, dx: float = 100.0, seed: int = 20261203) -> dict:
    """A coarse basin for the sub-grid channel path: n x n cells of dx (100 m), fBm
    (H 0.8, 800 m relief), pit-filled; rivers wherever the D8 drained area reaches
    1 km^2, carried as sub-grid channels (w = 8 A^0.5, d = 0.27 A^0.3, channel n
    0.03); floodplain Manning in [0.04, 0.10]; lognormal rain (median 15 mm/h, sigma
    0.6, 20 km correlation); infiltration U[0, 10] mm/h.  Rates in m/s."""
    dem = fbm_dem(n, n, hurst=0.8, relief=800.0, seed=seed)
    bed, acc = d8_area(dem, dx)
    width, depth = hydraulic_geometry(acc.astype(np.float64) * dx * dx / 1.0e6, dx)
    manning = voronoi_field(n, n, sites=512, lo=0.04, hi=0.10, seed=seed + 1)
    rain = lognormal_field(n, n, median=15.0, sigma=0.6, corr_cells=20000.0 / dx,
                           seed=seed + 2) * np.float32(MM_H)
    infil = uniform_field(n, n, 0.0, 10.0, seed=seed + 3) * np.float32(MM_H)
    return {"bed": bed, "manning": manning, "rain": rain.astype(np.float32),
            "infil": infil.astype(np.float32), "dx": dx, "width": width, "depth": depth,
            "chan_manning": np.full((n, n), 0.03, np.float32)}


def valley(n: int, *, dx: float = 5.0, seed: int = 20261204) -> dict:
    """The HLLC valley: an n x n grid of dx cells, a trapezoidal main channel 40 m wide
    and 3 m deep with 1:2 banks down the middle column, the bed falling 5e-4 down the
    rows (north to south), and 2 m of fBm relief on the floodplain clear of the channel.
    Returns the bed and the channel's columns (where the inflow enters on row 0)."""
    rows = np.arange(n, dtype=np.float64)[:, None]
    cols = np.arange(n, dtype=np.float64)[None, :]
    centre = n / 2.0
    off = np.abs(cols - centre) * dx  # metres from the channel axis
    channel = np.clip(3.0 - np.maximum(off - 20.0, 0.0) * 0.5, 0.0, 3.0)
    relief = fbm_dem(n, n, hurst=0.8, relief=2.0, seed=seed)
    bed = 5e-4 * dx * (n - rows) - channel + 3.0 + relief * (off > 26.0)
    inflow_cols = [int(j) for j in range(n) if abs(j - centre) * dx <= 20.0]
    return {"bed": bed.astype(np.float32), "channel_cols": inflow_cols, "dx": dx}
