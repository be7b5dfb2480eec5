"""Python face of the CPU restatement oracle (TEST INFRASTRUCTURE ONLY).

``liborc.so`` (``rbs_oracle.c``) restates the reference's Warp kernels pass by pass
in float32; this module holds the numpy field container that feeds it and the
host-side pieces of the reference that are plain Python/numpy there too:

* ``stoker`` / ``ritter`` / ``stoker_front_position`` -- validation/analytical.py:27-127
* ``MassLedger`` -- solver/core/massbalance.py:39-158
* ``stage_at`` -- solver/core/boundaries.py:49-70
* ``discharge_at`` -- solver/io/config.py:111-121

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
import this module.  It is the checker, never the thing measured or shipped.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB_PATH = _HERE / "liborc.so"

GRAVITY = 9.81
H_DRY = 1.0e-3
MASS_GATE = 1.0e-6
EDGES = ("west", "east", "north", "south")
_EDGE_TYPES = {"closed": 0, "open": 1, "fixed_stage": 2}

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class _OrcState(ctypes.Structure):
    _fields_ = [
        ("ny", ctypes.c_int32),
        ("nx", ctypes.c_int32),
        ("dx", ctypes.c_double),
        ("h", _fp), ("z", _fp), ("n", _fp), ("qx", _fp), ("qy", _fp),
        ("eta", _fp), ("beta", _fp),
        ("h_comp", _fp), ("rain", _fp), ("infil", _fp), ("loss_cum", _dp),
        ("bc", ctypes.c_int32 * 4),
        ("hu", _fp), ("hv", _fp), ("uvel", _fp), ("vvel", _fp),
        ("fx_h", _fp), ("fx_mn_l", _fp), ("fx_mn_r", _fp), ("fx_mt", _fp),
        ("fy_h", _fp), ("fy_mn_l", _fp), ("fy_mn_r", _fp), ("fy_mt", _fp),
        ("dh", _fp), ("dhu", _fp), ("dhv", _fp), ("h1", _fp), ("hu1", _fp), ("hv1", _fp),
        ("d50", _fp), ("qs_int_x", _fp), ("qs_comp_x", _fp),
        ("qs_int_y", _fp), ("qs_comp_y", _fp),
        ("ch_w", _fp), ("ch_d", _fp), ("ch_n", _fp),
        ("qx_ch", _fp), ("qx_fp", _fp), ("qy_ch", _fp), ("qy_fp", _fp),
    ]  # fmt: skip


_lib = None


def lib() -> ctypes.CDLL:
    """Load ``liborc.so`` (built by ``make -C oracle``)."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(f"oracle library {_LIB_PATH} is not built (make -C oracle)")
        L = ctypes.CDLL(str(_LIB_PATH))
        sp = ctypes.POINTER(_OrcState)
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_li_reduce_hmax.argtypes = [sp]
        L.orc_li_reduce_hmax.restype = ctypes.c_float
        L.orc_li_compute_dt.argtypes = [sp, ctypes.c_double, ctypes.c_double]
        L.orc_li_compute_dt.restype = ctypes.c_double
        L.orc_li_step.argtypes = [sp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                  ctypes.c_double]
        L.orc_li_run_fixed.argtypes = [sp, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_double]
        L.orc_hllc_reduce_wavespeed.argtypes = [sp]
        L.orc_hllc_reduce_wavespeed.restype = ctypes.c_float
        L.orc_hllc_compute_dt.argtypes = [sp, ctypes.c_double, ctypes.c_double]
        L.orc_hllc_compute_dt.restype = ctypes.c_double
        L.orc_hllc_step.argtypes = [sp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _fp]
        L.orc_hllc_flux.argtypes = [ctypes.c_float] * 6 + [_fp]
        L.orc_face_capacity.argtypes = [ctypes.c_float] * 5
        L.orc_face_capacity.restype = ctypes.c_float
        L.orc_shields_number.argtypes = [ctypes.c_float] * 5
        L.orc_shields_number.restype = ctypes.c_float
        L.orc_accumulate_transport.argtypes = [sp, ctypes.c_double]
        L.orc_exner_update.argtypes = [ctypes.c_int32, ctypes.c_int32, _fp, _fp, _fp, _fp,
                                       ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_rebuild_bed.argtypes = [ctypes.c_int32, ctypes.c_int32, _fp, _dp, _fp]
        L.orc_inject_point_sources_c.argtypes = [sp, ctypes.c_int32, _ip, _ip, _fp, _fp]
        for fn in ("orc_eta_subgrid", "orc_column_depth"):
            getattr(L, fn).argtypes = [ctypes.c_float] * (5 if fn == "orc_eta_subgrid" else 4)
            getattr(L, fn).restype = ctypes.c_float
        L.orc_manning_denominator_radius.argtypes = [ctypes.c_float] * 6
        L.orc_manning_denominator_radius.restype = ctypes.c_float
        for fn in ("orc_pow73", "orc_pow43", "orc_cbrt", "orc_pow15"):
            getattr(L, fn).argtypes = [ctypes.c_float]
            getattr(L, fn).restype = ctypes.c_float
        L.orc_ln_coarse.argtypes = [ctypes.c_float]
        L.orc_ln_coarse.restype = ctypes.c_double
        L.orc_set_pow_exact.argtypes = [ctypes.c_int]
        L.orc_get_pow_exact.restype = ctypes.c_int
        L.orc_pow_sweep.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
        L.orc_pow_sweep.restype = ctypes.c_long
        L.orc_pow_misses.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32,
                                     ctypes.POINTER(ctypes.c_uint32), ctypes.c_long]
        L.orc_pow_misses.restype = ctypes.c_long
        L.orc_fill_span.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_fill_span.restype = ctypes.c_double
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def set_pow(mode: str) -> None:
    """'exact': the correctly rounded powers with the reference's float32 exponents
    (the GPU's sequence, the default); 'libm': glibc powf with the same exponents."""
    if mode not in ("exact", "libm"):
        raise ValueError(mode)
    lib().orc_set_pow_exact(1 if mode == "exact" else 0)


def get_pow() -> str:
    return "exact" if lib().orc_get_pow_exact() else "libm"


class pow_mode:
    """Context manager: run the oracle with the given power mode, then restore."""

    def __init__(self, mode: str):
        self.mode = mode

    def __enter__(self):
        self.prev = get_pow()
        set_pow(self.mode)
        return self

    def __exit__(self, *exc):
        set_pow(self.prev)
        return False


def _ptr(a: np.ndarray | None, kind=_fp):
    if a is None:
        return ctypes.cast(None, kind)
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(kind)


def _f32(a, shape=None) -> np.ndarray:
    out = np.ascontiguousarray(np.array(a, dtype=np.float32, copy=True))
    if shape is not None and out.shape != shape:
        raise ValueError(f"shape {out.shape} != {shape}")
    return out


class State:
    """numpy field container with the reference ``State``'s fields and arming rules.

    Mirrors solver/core/state.py:38-226: float32 fields, optional fields armed on
    demand, ``loss_cum`` float64, all-closed edges by default.
    """

    def __init__(self, bed, dx, *, depth=0.0, manning=0.035):
        bed = _f32(bed)
        ny, nx = bed.shape
        self.ny, self.nx, self.dx = ny, nx, float(dx)
        self.z = bed
        self.h = (np.full((ny, nx), float(depth), np.float32) if np.isscalar(depth)
                  else _f32(depth, (ny, nx)))
        self.n = (np.full((ny, nx), float(manning), np.float32) if np.isscalar(manning)
                  else _f32(manning, (ny, nx)))
        self.qx = np.zeros((ny, nx + 1), np.float32)
        self.qy = np.zeros((ny + 1, nx), np.float32)
        self.eta = (self.h + self.z).astype(np.float32)
        self.beta = np.zeros((ny, nx), np.float32)
        self.h_comp = None
        self.rain = None
        self.infil = None
        self.loss_cum = None
        self.boundaries = {e: "closed" for e in EDGES}
        self.stage_curves: dict[str, list[tuple[float, float]]] = {}
        self.hu = self.hv = None
        self._hl = None
        self.d50 = None
        self.qs_int_x = self.qs_comp_x = self.qs_int_y = self.qs_comp_y = None
        self.ch_w = self.ch_d = self.ch_n = None
        self.qx_ch = self.qx_fp = self.qy_ch = self.qy_fp = None

    # --- arming (state.py:98-171) ---
    def set_infiltration(self, infil):
        self.infil = _f32(infil, (self.ny, self.nx))
        self._ensure_loss_cum()

    def set_rain_field(self, rain):
        self.rain = _f32(rain, (self.ny, self.nx))

    def set_open_boundaries(self, boundaries, stage_curves=None):
        self.boundaries = {e: boundaries.get(e, "closed") for e in EDGES}
        self.stage_curves = {e: list(c) for e, c in (stage_curves or {}).items()
                             if self.boundaries.get(e) == "fixed_stage"}
        missing = [e for e, v in self.boundaries.items()
                   if v == "fixed_stage" and e not in self.stage_curves]
        if missing:
            raise ValueError(f"fixed_stage edges without a stage curve: {sorted(missing)}")
        if any(v != "closed" for v in self.boundaries.values()):
            self._ensure_loss_cum()

    def arm_source_compensation(self):
        if self.h_comp is None:
            self.h_comp = np.zeros((self.ny, self.nx), np.float32)

    def _ensure_loss_cum(self):
        if self.loss_cum is None:
            self.loss_cum = np.zeros((self.ny, self.nx), np.float64)

    def arm_hllc(self):
        """hllc.py:123-152 (momentum starts at rest, scratch zeroed)."""
        if self.hu is not None:
            return
        ny, nx = self.ny, self.nx
        self.hu = np.zeros((ny, nx), np.float32)
        self.hv = np.zeros((ny, nx), np.float32)
        c = lambda: np.zeros((ny, nx), np.float32)  # noqa: E731
        xf = lambda: np.zeros((ny, nx + 1), np.float32)  # noqa: E731
        yf = lambda: np.zeros((ny + 1, nx), np.float32)  # noqa: E731
        self._hl = dict(uvel=c(), vvel=c(), fx_h=xf(), fx_mn_l=xf(), fx_mn_r=xf(), fx_mt=xf(),
                        fy_h=yf(), fy_mn_l=yf(), fy_mn_r=yf(), fy_mt=yf(), dh=c(), dhu=c(),
                        dhv=c(), h1=c(), hu1=c(), hv1=c())

    def arm_sediment(self, d50):
        """sediment.py:658-709 (face integrals and their Kahan debt start at zero)."""
        ny, nx = self.ny, self.nx
        self.d50 = (np.full((ny, nx), float(d50), np.float32) if np.isscalar(d50)
                    else _f32(d50, (ny, nx)))
        self.qs_int_x = np.zeros((ny, nx + 1), np.float32)
        self.qs_comp_x = np.zeros((ny, nx + 1), np.float32)
        self.qs_int_y = np.zeros((ny + 1, nx), np.float32)
        self.qs_comp_y = np.zeros((ny + 1, nx), np.float32)

    def arm_channels(self, width, depth, manning=None):
        """channels.py:289-330 with validate_geometry's normalisation (:275-277):
        a cell without both a width and a depth carries w = d = 0; the roughness
        defaults to the floodplain's; the face components start at zero."""
        ny, nx = self.ny, self.nx
        w = _f32(width, (ny, nx))
        d = _f32(depth, (ny, nx))
        n = self.n.copy() if manning is None else _f32(manning, (ny, nx))
        has = (w > 0) & (d > 0)
        w[~has] = 0.0
        d[~has] = 0.0
        self.ch_w, self.ch_d, self.ch_n = w, d, n
        self.qx_ch = np.zeros((ny, nx + 1), np.float32)
        self.qx_fp = np.zeros((ny, nx + 1), np.float32)
        self.qy_ch = np.zeros((ny + 1, nx), np.float32)
        self.qy_fp = np.zeros((ny + 1, nx), np.float32)

    def loss_volume(self, cell_area: float) -> float:
        if self.loss_cum is None:
            return 0.0
        return float(self.loss_cum.astype(np.float64).sum()) * cell_area

    @property
    def cell_area(self) -> float:
        return self.dx * self.dx

    def _struct(self) -> _OrcState:
        s = _OrcState()
        s.ny, s.nx, s.dx = self.ny, self.nx, self.dx
        for name in ("h", "z", "n", "qx", "qy", "eta", "beta", "h_comp", "rain", "infil",
                     "hu", "hv", "d50", "qs_int_x", "qs_comp_x", "qs_int_y", "qs_comp_y",
                     "ch_w", "ch_d", "ch_n", "qx_ch", "qx_fp", "qy_ch", "qy_fp"):
            setattr(s, name, _ptr(getattr(self, name)))
        s.loss_cum = _ptr(self.loss_cum, _dp)
        for k, e in enumerate(EDGES):
            s.bc[k] = _EDGE_TYPES[self.boundaries[e]]
        if self._hl is not None:
            for name, arr in self._hl.items():
                setattr(s, name, _ptr(arr))
        self._keep = s
        return s


# --- the scheme seam (schemes.py:1-52): compute_dt / step ----------------------
def li_compute_dt(st: State, alpha: float = 0.7, dt_max: float = 30.0) -> float:
    return float(lib().orc_li_compute_dt(ctypes.byref(st._struct()), alpha, dt_max))


def li_step(st: State, dt: float, rain: float = 0.0, limit: bool = True,
            rain_scale: float = 1.0, t: float = 0.0) -> None:
    lib().orc_li_step(ctypes.byref(st._struct()), float(dt), float(rain), int(bool(limit)),
                      float(rain_scale))


def li_run_fixed(st: State, dt: float, steps: int, rain: float = 0.0,
                 rain_scale: float = 1.0) -> None:
    lib().orc_li_run_fixed(ctypes.byref(st._struct()), float(dt), int(steps), float(rain),
                           float(rain_scale))


def hllc_compute_dt(st: State, alpha: float = 0.45, dt_max: float = 30.0) -> float:
    st.arm_hllc()
    return float(lib().orc_hllc_compute_dt(ctypes.byref(st._struct()), alpha, dt_max))


def hllc_step(st: State, dt: float, rain: float = 0.0, rain_scale: float = 1.0,
              t: float = 0.0) -> None:
    """hllc.py:1272-1373; stages evaluated on the host at t + dt/2 (hllc.py:1308-1312)."""
    st.arm_hllc()
    stages = np.zeros(4, np.float32)
    for k, e in enumerate(EDGES):
        if e in st.stage_curves:
            stages[k] = np.float32(stage_at(st.stage_curves[e], t + 0.5 * float(dt)))
    lib().orc_hllc_step(ctypes.byref(st._struct()), float(dt), float(rain), float(rain_scale),
                        _ptr(stages))


def hllc_flux(hl, ul, vtl, hr, ur, vtr) -> tuple[float, float, float]:
    out = np.zeros(3, np.float32)
    lib().orc_hllc_flux(hl, ul, vtl, hr, ur, vtr, _ptr(out))
    return float(out[0]), float(out[1]), float(out[2])


def face_capacity(q, h, radius, n, d50) -> float:
    return float(lib().orc_face_capacity(q, h, radius, n, d50))


def shields_number(q, h, radius, n, d50) -> float:
    return float(lib().orc_shields_number(q, h, radius, n, d50))


def accumulate_transport(st: State, dt: float) -> None:
    lib().orc_accumulate_transport(ctypes.byref(st._struct()), float(dt))


def exner_update(qs_int_x, qs_int_y, dz_lo, dz_hi, dx, porosity, dz_cum, dz_unapplied):
    ny, nx = dz_cum.shape
    a = [_f32(qs_int_x, (ny, nx + 1)), _f32(qs_int_y, (ny + 1, nx)),
         _f32(dz_lo, (ny, nx)), _f32(dz_hi, (ny, nx))]
    assert dz_cum.dtype == np.float64 and dz_unapplied.dtype == np.float64
    lib().orc_exner_update(ny, nx, *(_ptr(x) for x in a), float(dx),
                           1.0 / (1.0 - float(porosity)), _ptr(dz_cum, _dp),
                           _ptr(dz_unapplied, _dp))


def rebuild_bed(z0, dz_cum, z):
    ny, nx = z.shape
    lib().orc_rebuild_bed(ny, nx, _ptr(np.ascontiguousarray(z0, np.float32)), _ptr(dz_cum, _dp),
                          _ptr(z))


def inject_point_sources_c(st: State, rows, cols, add_depth, comp):
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    add = np.ascontiguousarray(add_depth, np.float32)
    lib().orc_inject_point_sources_c(ctypes.byref(st._struct()), len(rows), _ptr(rows, _ip),
                                     _ptr(cols, _ip), _ptr(add), _ptr(comp))


def pow73(x: float) -> float:
    return float(lib().orc_pow73(x))


def ln_coarse(x: float) -> float:
    return float(lib().orc_ln_coarse(x))


def pow_sweep(kind: int, lo: int, hi: int, libm: bool = False) -> int:
    """Float32 bit patterns in [lo, hi) whose power (kind 0..3: x^f32(7/3), x^f32(4/3),
    x^f32(1/3), x^1.5) differs from powl's with the same exponent, rounded; libm
    counts glibc powf's instead of the oracle's sequence."""
    return int(lib().orc_pow_sweep(kind, lo, hi, 1 if libm else 0))


def pow_misses(kind: int, lo: int, hi: int, maxn: int = 64) -> np.ndarray:
    """The float32 inputs in [lo, hi) (bit patterns) where the oracle's power differs
    from powl's, rounded (kind 0..2), at most maxn of them."""
    out = np.zeros(maxn, dtype=np.uint32)
    n = lib().orc_pow_misses(kind, lo, hi, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), maxn)
    return out[:n].view(np.float32).copy()


def pow43(x: float) -> float:
    return float(lib().orc_pow43(x))


def eta_subgrid(h: float, z: float, w: float, d: float, dx: float) -> float:
    return float(lib().orc_eta_subgrid(h, z, w, d, dx))


def column_depth(h: float, w: float, d: float, dx: float) -> float:
    return float(lib().orc_column_depth(h, w, d, dx))


def manning_denominator_radius(q, h_flow, radius, n, g, dt) -> float:
    return float(lib().orc_manning_denominator_radius(q, h_flow, radius, n, g, dt))


def cbrt(x: float) -> float:
    return float(lib().orc_cbrt(x))


def pow15(x: float) -> float:
    return float(lib().orc_pow15(x))


def fill_span(span: float, dt_raw: float) -> float:
    return float(lib().orc_fill_span(span, dt_raw))


# --- reservoirs (processes/reservoir.py:83-270, io/config.py:239-255) --------------------
def release_rule(rule: dict, stage):
    """Structure.discharge_at: ``rule`` holds release_rule and its parameters."""
    if rule["release_rule"] == "none" or stage is None:
        return 0.0
    if rule["release_rule"] == "fixed":
        return float(rule["release_m3_s"])
    span = rule["crest_m"] - float(rule["target_stage_m"])
    frac = (stage - float(rule["target_stage_m"])) / span
    return float(rule["release_max_m3_s"]) * min(max(frac, 0.0), 1.0)


def reservoir_release(st: State, rule: dict, pool, outlet, gauge, t: float, dt_slow: float):
    """ReservoirOperator.advance over an oracle state: the gauge stage, the rule, then
    _withdraw_pool and _deliver_release in float32 with float64 sums, the delivery's
    rounding banked into loss_cum.  Returns (time, stage, discharge, volume)."""
    st._ensure_loss_cum()
    gi, gj = gauge
    h = float(st.h[gi, gj])
    stage = None if h <= H_DRY else float(st.z[gi, gj]) + h
    q = release_rule(rule, stage)
    volume = 0.0
    if q > 0.0:
        area = st.cell_area
        r0, c0, r1, c1 = pool
        ps = (slice(r0, r1 + 1), slice(c0, c1 + 1))
        stored = float(st.h[ps].astype(np.float64).sum()) * area
        if stored > 0.0:
            scale = np.float32(min(1.0, (q * float(dt_slow)) / stored))
            avail = st.h[ps].copy()
            h_new = avail - avail * scale
            st.h[ps] = h_new
            removed = float((avail.astype(np.float64) - h_new.astype(np.float64)).sum()) * area
            if removed > 0.0:
                o0, p0, o1, p1 = outlet
                os_ = (slice(o0, o1 + 1), slice(p0, p1 + 1))
                n_out = (o1 - o0 + 1) * (p1 - p0 + 1)
                add = removed / area / n_out
                av = st.h[os_].copy()
                hn = av + np.float32(add)
                st.h[os_] = hn
                st.loss_cum[os_] += add - (hn.astype(np.float64) - av.astype(np.float64))
                volume = removed
    return float(t), stage, q, volume


# --- host-side reference pieces ---------------------------------------------------
def stage_at(curve, t: float) -> float:
    """boundaries.py:49-70 -- piecewise linear, held outside its range."""
    if not curve:
        raise ValueError("empty stage curve")
    if t <= curve[0][0]:
        return float(curve[0][1])
    if t >= curve[-1][0]:
        return float(curve[-1][1])
    for (t0, s0), (t1, s1) in zip(curve, curve[1:]):
        if t0 <= t <= t1:
            if t1 == t0:
                return float(s1)
            return float(s0 + (s1 - s0) * (t - t0) / (t1 - t0))
    return float(curve[-1][1])


def discharge_at(hydrograph, t: float) -> float:
    """config.py:111-121 -- piecewise linear, zero outside its range."""
    hg = hydrograph
    if not hg or t < hg[0][0] or t > hg[-1][0]:
        return 0.0
    for (t0, q0), (t1, q1) in zip(hg, hg[1:]):
        if t0 <= t <= t1:
            if t1 == t0:
                return q1
            return q0 + (q1 - q0) * (t - t0) / (t1 - t0)
    return hg[-1][1]


@dataclass
class _Kahan:
    """massbalance.py:39-50."""

    total: float = 0.0
    comp: float = 0.0

    def add(self, x: float) -> None:
        y = x - self.comp
        t = self.total + y
        self.comp = (t - self.total) - y
        self.total = t


@dataclass
class MassLedger:
    """massbalance.py:65-158 over an oracle ``State``."""

    cell_area: float
    v0: float
    _inflow: _Kahan = field(default_factory=_Kahan)
    _outflow: _Kahan = field(default_factory=_Kahan)
    series: list = field(default_factory=list)
    _peak_v: float = 0.0

    @classmethod
    def from_state(cls, st: State) -> MassLedger:
        v0 = float(st.h.astype(np.float64).sum()) * st.cell_area
        led = cls(cell_area=st.cell_area, v0=v0)
        led.record(st, 0.0)
        return led

    def add_inflow(self, volume: float) -> None:
        self._inflow.add(volume)

    def add_outflow(self, volume: float) -> None:
        self._outflow.add(volume)

    def add_rain_step(self, rain_m_s: float, dt: float, n_cells: int) -> None:
        self.add_inflow(rain_m_s * dt * self.cell_area * n_cells)

    def record(self, st: State, time: float) -> dict:
        v = float(st.h.astype(np.float64).sum()) * self.cell_area
        inflow = self._inflow.total
        outflow = self._outflow.total + st.loss_volume(self.cell_area)
        residual = inflow - outflow - (v - self.v0)
        self._peak_v = max(self._peak_v, v)
        denom = max(abs(inflow), abs(v), self._peak_v, 1e-12)
        rec = dict(time=time, volume=v, inflow_cum=inflow, outflow_cum=outflow,
                   residual=residual, rel_error=abs(residual) / denom)
        self.series.append(rec)
        return rec

    @property
    def max_rel_error(self) -> float:
        return max((r["rel_error"] for r in self.series), default=0.0)


def stoker_star_depth(h_l: float, h_r: float, g: float = GRAVITY, tol: float = 1e-12) -> float:
    """analytical.py:27-53 -- bisection for the wet-bed star depth."""

    def f(h_m):
        u_rare = 2.0 * (math.sqrt(g * h_l) - math.sqrt(g * h_m))
        u_shock = (h_m - h_r) * math.sqrt(0.5 * g * (h_m + h_r) / (h_m * h_r))
        return u_rare - u_shock

    lo, hi = h_r, h_l
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(mid) > 0.0:
            lo = mid
        else:
            hi = mid
        if hi - lo < tol:
            break
    return 0.5 * (lo + hi)


def stoker(x, t: float, h_l: float, h_r: float, g: float = GRAVITY):
    """analytical.py:56-92 -- wet-bed dam-break (h, u) at x (dam at 0)."""
    x = np.asarray(x, dtype=np.float64)
    if t <= 0.0:
        return np.where(x < 0.0, h_l, h_r), np.zeros_like(x)
    c_l = math.sqrt(g * h_l)
    h_m = stoker_star_depth(h_l, h_r, g)
    c_m = math.sqrt(g * h_m)
    u_m = 2.0 * (c_l - c_m)
    s = h_m * u_m / (h_m - h_r)
    xi = x / t
    fan_tail = u_m - c_m
    h = np.empty_like(x)
    u = np.empty_like(x)
    left = xi <= -c_l
    fan = (xi > -c_l) & (xi < fan_tail)
    star = (xi >= fan_tail) & (xi < s)
    right = xi >= s
    h[left], u[left] = h_l, 0.0
    c_fan = (1.0 / 3.0) * (2.0 * c_l - xi[fan])
    h[fan], u[fan] = c_fan * c_fan / g, (2.0 / 3.0) * (xi[fan] + c_l)
    h[star], u[star] = h_m, u_m
    h[right], u[right] = h_r, 0.0
    return h, u


def ritter(x, t: float, h_l: float, g: float = GRAVITY):
    """analytical.py:95-117 -- dry-bed dam-break (h, u)."""
    x = np.asarray(x, dtype=np.float64)
    if t <= 0.0:
        return np.where(x < 0.0, h_l, 0.0), np.zeros_like(x)
    c_l = math.sqrt(g * h_l)
    xi = x / t
    h = np.empty_like(x)
    u = np.empty_like(x)
    left = xi <= -c_l
    fan = (xi > -c_l) & (xi < 2.0 * c_l)
    dry = xi >= 2.0 * c_l
    h[left], u[left] = h_l, 0.0
    h[fan] = (1.0 / (9.0 * g)) * (2.0 * c_l - xi[fan]) ** 2
    u[fan] = (2.0 / 3.0) * (c_l + xi[fan])
    h[dry], u[dry] = 0.0, 0.0
    return h, u


def stoker_front_position(t: float, h_l: float, h_r: float, g: float = GRAVITY) -> float:
    """analytical.py:120-127."""
    h_m = stoker_star_depth(h_l, h_r, g)
    c_l = math.sqrt(g * h_l)
    c_m = math.sqrt(g * h_m)
    u_m = 2.0 * (c_l - c_m)
    return h_m * u_m / (h_m - h_r) * t
