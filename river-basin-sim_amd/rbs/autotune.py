"""LI kernel-variant autotuner.

librbs carries several fused local-inertial kernels (LDS tiles of several shapes,
register strips of several heights, and strips side by side in row blocks whose stores
are whole cache lines); all of them write the same bits, and which is fastest depends
on the grid shape (a small grid needs small work units to fill 256 CUs, a wide one
gains from line-aligned stores) and on which sources are armed.  ``best_li_variant`` times every variant on a
synthetic state of the given shape and arming (HIP events around each fused kernel,
``rbs_time_li_kernel``), keeps the fastest, and caches the choice per key for the
life of the process.  ``tune`` applies it to a state.
"""

from __future__ import annotations

import ctypes

import numpy as np

from rbs import _lib

_cache: dict[tuple, int] = {}


def variant_table() -> list[dict]:
    """Every variant: index, family, tile shape, and whether it is the library default."""
    count = ctypes.c_int32()
    _lib.call("rbs_li_variant_count", ctypes.byref(count))
    out = []
    for v in range(count.value):
        bx, by, fam, dflt = (ctypes.c_int32() for _ in range(4))
        _lib.call("rbs_li_variant_info", v, ctypes.byref(bx), ctypes.byref(by), ctypes.byref(fam),
                  ctypes.byref(dflt))
        out.append({"variant": v, "family": ("lds_tile", "strip", "strip2", "rows")[fam.value],
                    "tile": (bx.value, by.value), "default": bool(dflt.value)})
    return out


def _probe_state(ny: int, nx: int, armed: int):
    from rbs.core.sediment import arm_sediment
    from rbs.core.state import State

    rng = np.random.default_rng(7)
    yy, xx = np.mgrid[0:ny, 0:nx]
    bed = (0.002 * (xx + yy) + 0.2 * rng.uniform(0.0, 1.0, (ny, nx))).astype(np.float32)
    st = State.from_bed(bed, 10.0, depth=0.3, manning=0.035)
    if armed & _lib.ARM_RAIN:
        st.set_rain_field(np.full((ny, nx), 5e-6, np.float32))
    if armed & _lib.ARM_INFIL:
        st.set_infiltration(np.full((ny, nx), 1e-6, np.float32))
    if armed & _lib.ARM_COMP:
        st.arm_source_compensation()
    if armed & _lib.ARM_SED:
        arm_sediment(st, np.full((ny, nx), 1e-3, np.float32), 0.4)
    st.set_open_boundaries({e: "open" for e in ("north", "south", "east", "west")})
    return st


def _armed(state) -> int:
    flags = ctypes.c_uint32()
    _lib.call("rbs_armed", state._handle, ctypes.byref(flags))
    return int(flags.value)


def _bucket(n: int) -> int:
    """Grid sides are tuned in buckets of 128 cells (the probe uses the bucket's size)."""
    return max(128, -(-int(n) // 128) * 128)


def best_li_variant(ny: int, nx: int, armed: int = 0, reps: int = 12) -> int:
    """The fastest LI variant for a (ny, nx) grid with the given arm flags."""
    key = (_bucket(ny), _bucket(nx), int(armed) & (_lib.ARM_RAIN | _lib.ARM_INFIL
                                                   | _lib.ARM_COMP | _lib.ARM_SED))
    if key in _cache:
        return _cache[key]
    st = _probe_state(key[0], key[1], key[2])
    try:
        best, best_ms = -1, float("inf")
        for row in variant_table():
            _lib.call("rbs_set_li_variant", st._handle, row["variant"])
            avg = ctypes.c_double()
            # one untimed pass settles the caches, then the timed reps
            _lib.call("rbs_time_li_kernel", st._handle, 2, 0.2, 0.0, 1.0, ctypes.byref(avg))
            _lib.call("rbs_time_li_kernel", st._handle, reps, 0.2, 0.0, 1.0, ctypes.byref(avg))
            if avg.value < best_ms:
                best, best_ms = row["variant"], avg.value
    finally:
        st.close()
    _cache[key] = best
    return best


def tune(state) -> int:
    """Switch ``state`` to the fastest LI variant for its shape and arming."""
    ny, nx = state.grid.shape
    v = best_li_variant(ny, nx, _armed(state))
    _lib.call("rbs_set_li_variant", state._handle, v)
    return v
