"""Error behaviour at the C-ABI, as the reference raises it.

The reference raises ValueError for bad shapes, an unknown scheme or a fixed_stage edge
without a curve (schemes.py:38-52, state.py:113-143) and RuntimeError for a
non-positive timestep (scheduler.py:221-227).  librbs returns RBS_E_ARG and
RBS_E_RANGE for those classes and `rbs._lib` maps them to the same exceptions; these
tests call the entry points directly, below the Python checks, so the library's own
refusal is what is tested.  Every refused call leaves the state usable.
"""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rbs():
    from rbs import _lib
    from rbs.core import local_inertial, schemes, state

    _lib.load()
    return type("R", (), dict(lib=_lib, li=local_inertial, schemes=schemes, State=state.State))


def _state(rbs, ny=8, nx=12):
    return rbs.State.from_bed(np.zeros((ny, nx), np.float32), 2.0, depth=0.5, manning=0.03)


@pytest.mark.parametrize("ny,nx,dx", [(0, 8, 1.0), (8, 0, 1.0), (0, 0, 1.0), (4, 4, 0.0),
                                      (4, 4, -1.0), (4, 4, float("nan"))])
def test_an_empty_or_degenerate_grid_is_refused(rbs, ny, nx, dx):
    handle = ctypes.c_void_p()
    lib = rbs.lib.load()
    rc = lib.rbs_state_create(ny, nx, ctypes.c_double(dx), 0, ctypes.byref(handle))
    assert rc == rbs.lib.E_ARG
    assert not handle.value
    assert "bad grid" in lib.rbs_last_error().decode()


@pytest.mark.parametrize("dt", [0.0, -1.0, float("nan"), float("-inf")])
def test_a_non_positive_timestep_is_refused_by_the_library(rbs, dt):
    st = _state(rbs)
    h0 = st.h.numpy().copy()
    with pytest.raises(RuntimeError):
        rbs.lib.call("rbs_li_step", st._handle, ctypes.c_double(dt), ctypes.c_double(0.0), 1,
                     ctypes.c_double(1.0), ctypes.c_double(0.0))
    with pytest.raises(RuntimeError):  # the Python seam refuses it before the library
        rbs.li.step(st, dt)
    assert np.array_equal(st.h.numpy(), h0)  # nothing ran
    rbs.li.step(st, 0.01)  # and the state still steps
    st.close()


def test_an_upload_of_the_wrong_size_is_refused(rbs):
    st = _state(rbs)
    wrong = np.zeros((8, 11), np.float32)
    with pytest.raises(ValueError):
        rbs.lib.call("rbs_upload", st._handle, rbs.lib.F_H, wrong.ctypes.data_as(ctypes.c_void_p),
                     ctypes.c_size_t(wrong.nbytes))
    with pytest.raises(ValueError):  # qx is (ny, nx + 1)
        rbs.lib.call("rbs_upload", st._handle, rbs.lib.F_QX,
                     np.zeros((8, 12), np.float32).ctypes.data_as(ctypes.c_void_p),
                     ctypes.c_size_t(8 * 12 * 4))
    assert np.all(st.h.numpy() == np.float32(0.5))
    st.close()


def test_boundaries_and_schemes_are_refused_as_the_reference_refuses_them(rbs):
    st = _state(rbs)
    with pytest.raises(ValueError, match="stage curve"):
        st.set_open_boundaries({"south": "fixed_stage"})
    with pytest.raises(ValueError):
        st.set_open_boundaries({"south": "weir"})
    with pytest.raises(ValueError, match="unknown scheme"):
        rbs.schemes.get_scheme("kinematic_wave")
    # a fixed_stage edge handed to the library without its curve is refused at the step
    types = (ctypes.c_int32 * 4)(0, 0, 0, rbs.lib.BC["fixed_stage"])
    rbs.lib.call("rbs_set_boundaries", st._handle, types)
    with pytest.raises(Exception, match="stage"):
        rbs.lib.call("rbs_hllc_step", st._handle, ctypes.c_double(0.01), ctypes.c_double(0.0),
                     ctypes.c_double(1.0), ctypes.c_double(0.0))
    st.close()
