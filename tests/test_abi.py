"""The C-ABI library loads and exports every entry point include/rbs.h declares.

CPU-only: nothing here touches a GPU (only rbs_version, a static string, is called).
"""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "rbs.h"
LIB = ROOT / "river-basin-sim_amd" / "rbs" / "librbs.so"


def declared_symbols() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rbs_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_seam():
    syms = declared_symbols()
    for s in ("rbs_state_create", "rbs_compute_dt", "rbs_li_step", "rbs_hllc_step",
              "rbs_upload", "rbs_download", "rbs_ledger_read", "rbs_run_until"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    if not LIB.exists():
        pytest.fail(f"{LIB} not built (python river-basin-sim_amd/build.py)")
    lib = ctypes.CDLL(str(LIB))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"librbs.so lacks {missing}"
    lib.rbs_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.rbs_version()


def test_binding_covers_the_header():
    """The ctypes binding declares a signature for every header symbol, and no more."""
    import sys

    sys.path.insert(0, str(ROOT / "river-basin-sim_amd"))
    from rbs import _lib

    assert sorted(_lib.SIGNATURES) == declared_symbols()


def test_library_is_built_for_gfx950_only():
    """The code object inside librbs.so targets gfx950 (no CUDA, no other arch)."""
    data = LIB.read_bytes()
    assert b"gfx950" in data
    assert b"sm_" + b"90" not in data
