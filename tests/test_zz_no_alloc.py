"""The no-allocation reservoir release (rbs_reservoir_release_noalloc) against the
allocating entry point it shares kernels with: the same depths, record and banked
rounding bit for bit, on two calls in a row, the second over a smaller pool (fewer
block partials than the first left in the shared scratch)."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _states(shape, pool_rows, seed):
    from rbs.core.state import State
    from rbs.processes.reservoir import ReservoirOperator, Structure

    ny, nx = shape
    rng = np.random.default_rng(seed)
    bed = rng.uniform(0.0, 0.5, shape).astype(np.float32)
    depth = np.zeros(shape, np.float32)
    depth[:pool_rows] = rng.uniform(0.5, 3.0, (pool_rows, nx)).astype(np.float32)
    dam = Structure(name="d", kind="dam", cells=[(pool_rows, j) for j in range(nx)], crest_m=4.0,
                    release_rule="fixed", release_m3_s=50.0, pool=(0, 0, pool_rows - 1, nx - 1),
                    outlet=(ny - 20, 10, ny - 5, 40), interval_s=300.0)
    a = State.from_bed(bed, 10.0, depth=depth, manning=0.03)
    b = State.from_bed(bed, 10.0, depth=depth, manning=0.03)
    return a, b, ReservoirOperator(dam, a), ReservoirOperator(dam, b)


def _release(name, st, op, t, dt_slow):
    from rbs import _lib

    out = _lib.Release()
    print(f"[no-alloc] {name} {st.grid.shape} pool rows {op.structure.pool[2] + 1}", flush=True)
    _lib.call(name, st._handle, ctypes.byref(op._params), float(t), float(dt_slow),
              ctypes.byref(out))
    return (out.time, out.stage, out.discharge_m3_s, out.volume_m3)


@pytest.mark.parametrize("shape,pool_rows,seed", [((2600, 64), 2500, 1), ((300, 64), 120, 2)])
def test_no_alloc_release_equals_the_allocating_one(shape, pool_rows, seed):
    from rbs import _lib

    _lib.load()
    a, b, op_a, op_b = _states(shape, pool_rows, seed)
    for k in range(2):
        t = 300.0 * (k + 1)
        ra = _release("rbs_reservoir_release", a, op_a, t, 300.0)
        rb = _release("rbs_reservoir_release_noalloc", b, op_b, t, 300.0)
        assert ra == rb
        assert ra[3] > 0.0
        assert np.array_equal(a.h.numpy(), b.h.numpy())
        assert a.ledger().release_out == b.ledger().release_out
    a.close()
    b.close()


def test_a_smaller_release_after_a_larger_one():
    """Two calls in a row on one GPU: a pool of 2500 rows (2048 block partials), then one
    of 90 rows; the second reads only its own partials."""
    from rbs import _lib

    _lib.load()
    big = _states((2600, 64), 2500, 3)
    small = _states((200, 64), 90, 4)
    for a, b, op_a, op_b in (big, small):
        ra = _release("rbs_reservoir_release", a, op_a, 300.0, 300.0)
        rb = _release("rbs_reservoir_release_noalloc", b, op_b, 300.0, 300.0)
        assert ra == rb and ra[3] > 0.0
        assert np.array_equal(a.h.numpy(), b.h.numpy())
        assert a.ledger().release_out == b.ledger().release_out
        a.close()
        b.close()
