"""ctypes binding of librbs.so (include/rbs.h).

The product path has no fallback: if the library is missing or no GPU is visible,
the first call raises.  Nothing here computes physics.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

_LIB_PATH = Path(os.environ.get("RBS_LIBRARY", Path(__file__).resolve().parent / "librbs.so"))

# error codes (rbs.h)
OK, E_ARG, E_STATE, E_HIP, E_NOMEM, E_RANGE = 0, -1, -2, -3, -4, -5

# field ids (rbs.h enum rbs_field)
F_H, F_Z, F_N, F_QX, F_QY, F_HU, F_HV, F_RAIN, F_INFIL, F_H_COMP = range(10)
F_D50, F_QS_INT_X, F_QS_COMP_X, F_QS_INT_Y, F_QS_COMP_Y, F_Z0 = range(10, 16)
F_DZ_CUM, F_DZ_UNAPPLIED, F_DZ_LO, F_DZ_HI, F_ETA = range(16, 21)
F_CH_W, F_CH_D, F_CH_N, F_QX_CH, F_QX_FP, F_QY_CH, F_QY_FP = range(21, 28)

ARM_COMP, ARM_RAIN, ARM_INFIL, ARM_HLLC, ARM_SED, ARM_CHAN = 1, 2, 4, 8, 16, 32
FRAME_DEPTH, FRAME_U, FRAME_V, FRAME_BED_CHANGE = 1, 2, 4, 8
EDGE = {"west": 0, "east": 1, "north": 2, "south": 3}
BC = {"closed": 0, "open": 1, "fixed_stage": 2}
SCHEME_LI, SCHEME_HLLC = 0, 1


class RbsError(RuntimeError):
    """A librbs call failed (HIP error, bad state)."""

    def __init__(self, code: int, message: str):
        super().__init__(f"librbs error {code}: {message}")
        self.code = code


class Ledger(ctypes.Structure):
    _fields_ = [
        ("volume", ctypes.c_double),
        ("rain_in", ctypes.c_double), ("rain_in_c", ctypes.c_double),
        ("point_in", ctypes.c_double), ("point_in_c", ctypes.c_double),
        ("infil_out", ctypes.c_double), ("infil_out_c", ctypes.c_double),
        ("edge_out", ctypes.c_double), ("edge_out_c", ctypes.c_double),
        ("t", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("steps", ctypes.c_int64),
        ("speed_max", ctypes.c_float),
        ("active", ctypes.c_int32),
        ("release_out", ctypes.c_double), ("release_out_c", ctypes.c_double),
    ]  # fmt: skip


class Reservoir(ctypes.Structure):
    _fields_ = [
        ("pool", ctypes.c_int32 * 4),
        ("outlet", ctypes.c_int32 * 4),
        ("gauge_row", ctypes.c_int32), ("gauge_col", ctypes.c_int32),
        ("rule", ctypes.c_int32), ("pad", ctypes.c_int32),
        ("release_m3_s", ctypes.c_double),
        ("target_stage_m", ctypes.c_double),
        ("crest_m", ctypes.c_double),
        ("release_max_m3_s", ctypes.c_double),
    ]  # fmt: skip


class Release(ctypes.Structure):
    _fields_ = [("time", ctypes.c_double), ("stage", ctypes.c_double),
                ("discharge_m3_s", ctypes.c_double), ("volume_m3", ctypes.c_double)]


RULE_FIXED, RULE_TARGET_STAGE = 1, 2


class RunParams(ctypes.Structure):
    _fields_ = [
        ("scheme", ctypes.c_int32),
        ("limit", ctypes.c_int32),
        ("alpha", ctypes.c_double),
        ("dt_max", ctypes.c_double),
        ("rain_m_s", ctypes.c_double),
        ("rain_duration", ctypes.c_double),
        ("rain_is_field", ctypes.c_int32),
        ("n_sync", ctypes.c_int32),
        ("sync_times", ctypes.POINTER(ctypes.c_double)),
        ("rain_field_sum", ctypes.c_double),
    ]


# Every symbol the header declares, with its ctypes signature.
_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_dbl = ctypes.c_double
_pi32 = ctypes.POINTER(ctypes.c_int32)
_pdbl = ctypes.POINTER(ctypes.c_double)
SIGNATURES: dict[str, tuple] = {
    "rbs_last_error": (ctypes.c_char_p, []),
    "rbs_version": (ctypes.c_char_p, []),
    "rbs_device_count": (ctypes.c_int, [_pi32]),
    "rbs_state_create": (ctypes.c_int, [_i32, _i32, _dbl, ctypes.c_uint32, ctypes.POINTER(_vp)]),
    "rbs_state_destroy": (ctypes.c_int, [_vp]),
    "rbs_arm": (ctypes.c_int, [_vp, ctypes.c_uint32]),
    "rbs_armed": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint32)]),
    "rbs_state_bytes": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    "rbs_state_footprint": (ctypes.c_int, [_i32, _i32, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64),
                                            ctypes.POINTER(ctypes.c_uint64)]),
    "rbs_upload": (ctypes.c_int, [_vp, _i32, _vp, ctypes.c_size_t]),
    "rbs_download": (ctypes.c_int, [_vp, _i32, _vp, ctypes.c_size_t]),
    "rbs_set_boundaries": (ctypes.c_int, [_vp, _pi32]),
    "rbs_set_stage_curve": (ctypes.c_int, [_vp, _i32, _i32, _pdbl, _pdbl]),
    "rbs_compute_dt": (ctypes.c_int, [_vp, _i32, _dbl, _dbl, _pdbl]),
    "rbs_li_step": (ctypes.c_int, [_vp, _dbl, _dbl, _i32, _dbl, _dbl]),
    "rbs_hllc_step": (ctypes.c_int, [_vp, _dbl, _dbl, _dbl, _dbl]),
    "rbs_set_inflows": (ctypes.c_int, [_vp, _i32, _pi32, _pi32, _pi32, _pdbl, _pdbl]),
    "rbs_inflow_apply": (ctypes.c_int, [_vp, _dbl, _dbl, _pdbl]),
    "rbs_exner_activate": (ctypes.c_int, [_vp, _dbl]),
    "rbs_ledger_read": (ctypes.c_int, [_vp, ctypes.POINTER(Ledger)]),
    "rbs_clock_reset": (ctypes.c_int, [_vp, _dbl]),
    "rbs_run_until": (ctypes.c_int, [_vp, ctypes.POINTER(RunParams), _dbl, ctypes.c_int64,
                                     ctypes.POINTER(ctypes.c_int64)]),
    "rbs_sync": (ctypes.c_int, [_vp]),
    "rbs_batch_create": (ctypes.c_int, [_i32, ctypes.POINTER(_vp)]),
    "rbs_batch_destroy": (ctypes.c_int, [_vp]),
    "rbs_batch_add": (ctypes.c_int, [_vp, _vp, ctypes.POINTER(RunParams), _dbl, _pi32]),
    "rbs_batch_remove": (ctypes.c_int, [_vp, _i32]),
    "rbs_batch_advance": (ctypes.c_int, [_vp, _i32, _pi32]),
    "rbs_batch_member_active": (ctypes.c_int, [_vp, _i32, _pi32]),
    "rbs_set_li_variant": (ctypes.c_int, [_vp, _i32]),
    "rbs_get_li_variant": (ctypes.c_int, [_vp, _pi32]),
    "rbs_li_variant_count": (ctypes.c_int, [_pi32]),
    "rbs_li_variant_info": (ctypes.c_int, [_i32, _pi32, _pi32, _pi32, _pi32]),
    "rbs_batch_set_variant": (ctypes.c_int, [_vp, _i32, _i32]),
    "rbs_batch_get_variant": (ctypes.c_int, [_vp, _pi32, _pi32]),
    "rbs_frames_create": (ctypes.c_int, [_vp, ctypes.c_uint32, _i32, ctypes.POINTER(_vp)]),
    "rbs_frames_destroy": (ctypes.c_int, [_vp]),
    "rbs_frames_capture": (ctypes.c_int, [_vp, _pi32]),
    "rbs_frames_ready": (ctypes.c_int, [_vp, _i32, _pi32]),
    "rbs_frames_wait": (ctypes.c_int, [_vp, _i32, _pdbl] + [ctypes.POINTER(ctypes.POINTER(ctypes.c_float))] * 4),
    "rbs_frames_release": (ctypes.c_int, [_vp, _i32]),
    "rbs_frames_bytes": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    "rbs_reservoir_release": (ctypes.c_int, [_vp, ctypes.POINTER(Reservoir), _dbl, _dbl,
                                              ctypes.POINTER(Release)]),
    "rbs_reservoir_release_noalloc": (ctypes.c_int, [_vp, ctypes.POINTER(Reservoir), _dbl, _dbl,
                                              ctypes.POINTER(Release)]),
    "rbs_set_chan_variant": (ctypes.c_int, [_vp, _i32]),
    "rbs_chan_variant_count": (ctypes.c_int, [_pi32]),
    "rbs_chan_variant_info": (ctypes.c_int, [_i32, _pi32, _pi32, _pi32]),
    "rbs_set_hllc_variant": (ctypes.c_int, [_vp, _i32]),
    "rbs_hllc_variant_count": (ctypes.c_int, [_pi32]),
    "rbs_hllc_variant_info": (ctypes.c_int, [_i32, _pi32, _pi32, _pi32]),
    "rbs_time_hllc_step": (ctypes.c_int, [_vp, _i32, _dbl, _pdbl]),
    "rbs_batch_launch": (ctypes.c_int, [_vp, _i32]),
    "rbs_batch_poll": (ctypes.c_int, [_vp, _pi32]),
    "rbs_batch_collect": (ctypes.c_int, [_vp, _pi32]),
    "rbs_sed_totals": (ctypes.c_int, [_vp, _pdbl]),
    "rbs_batch_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64), _pdbl,
                                       ctypes.POINTER(ctypes.c_int64), _pdbl]),
    "rbs_time_li_kernel": (ctypes.c_int, [_vp, _i32, _dbl, _dbl, _dbl, _pdbl]),
}

_lib: ctypes.CDLL | None = None


def library_path() -> Path:
    return _LIB_PATH


def load() -> ctypes.CDLL:
    """Load librbs.so and attach signatures; raises if it is not built."""
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(
                f"{_LIB_PATH} is not built; run `python river-basin-sim_amd/build.py` "
                "(the solver has no CPU fallback)"
            )
        L = ctypes.CDLL(str(_LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    """Raise the Python error the reference raises for this failure class."""
    if rc == OK:
        return
    msg = load().rbs_last_error().decode(errors="replace")
    if rc == E_ARG:
        raise ValueError(msg)
    if rc == E_RANGE:
        raise RuntimeError(msg)
    raise RbsError(rc, msg)


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))


def device_count() -> int:
    n = ctypes.c_int32(0)
    call("rbs_device_count", ctypes.byref(n))
    return int(n.value)
