"""Scheme registry / dispatch (reference solver/core/schemes.py:1-52).

The run loop only calls ``compute_dt(state, alpha=..., dt_max=...)`` and
``step(state, dt=..., rain=..., rain_scale=..., t=...)`` on the module
``get_scheme`` returns.
"""

from __future__ import annotations

from types import ModuleType

from rbs.core import hllc, local_inertial

KNOWN_SCHEMES = ("local_inertial", "hllc_fv")


def get_scheme(name: str) -> ModuleType:
    """Module providing ``compute_dt``/``step`` for scheme ``name`` (schemes.py:38-52)."""
    if name == "local_inertial":
        return local_inertial
    if name == "hllc_fv":
        return hllc
    raise ValueError(f"unknown scheme '{name}'; known schemes: {list(KNOWN_SCHEMES)}")
