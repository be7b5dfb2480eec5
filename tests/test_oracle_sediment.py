"""Pin the oracle's sediment restatement (sediment.py) with the reference's
known-answer tests and closed forms (solver/core/test_sediment.py)."""

from __future__ import annotations

import numpy as np
import pytest

G = 9.81


def test_capacity_is_hand_computable(orc):
    """test_sediment.py:174-186: theta = 0.27273, q_s = 2.4699e-3 at q=2, h=1, n=.03, d=8 mm."""
    assert orc.shields_number(2.0, 1.0, 1.0, 0.03, 0.008) == pytest.approx(0.27273, rel=1e-4)
    assert orc.face_capacity(2.0, 1.0, 1.0, 0.03, 0.008) == pytest.approx(2.4699e-3, rel=1e-4)
    assert orc.face_capacity(-2.0, 1.0, 1.0, 0.03, 0.008) == pytest.approx(-2.4699e-3, rel=1e-4)


def test_below_threshold_is_bit_exact_zero(orc):
    """test_sediment.py:188-214."""
    q, n, d50 = 1.0, 0.03, 0.008
    h_c = (G * n**2 * q**2 / (1.65 * G * d50 * 0.047)) ** (3 / 7)
    for h in (h_c * 1.05, h_c * 2.0, 5.0):
        assert orc.face_capacity(q, h, h, n, d50) == 0.0
    for h in (h_c * 0.95, h_c * 0.5):
        assert orc.face_capacity(q, h, h, n, d50) > 0.0


def test_dry_face_zero_grain_zero_flow_carry_nothing(orc):
    assert orc.face_capacity(1.0, 0.5e-3, 0.5e-3, 0.03, 0.002) == 0.0
    assert orc.face_capacity(1.0, 0.5, 0.5, 0.03, 0.0) == 0.0
    assert orc.face_capacity(0.0, 0.5, 0.5, 0.03, 0.002) == 0.0


def _exner(orc, qs_x, qs_y, *, dx=10.0, porosity=0.4, lo=None, hi=None):
    ny, nx = qs_x.shape[0], qs_y.shape[1]
    lo = np.full((ny, nx), -np.inf, np.float32) if lo is None else lo
    hi = np.full((ny, nx), np.inf, np.float32) if hi is None else hi
    dz = np.zeros((ny, nx), np.float64)
    banked = np.zeros((ny, nx), np.float64)
    orc.exner_update(qs_x, qs_y, lo, hi, dx, porosity, dz, banked)
    return dz, banked


def test_uniform_transport_moves_no_bed(orc):
    """test_sediment.py:509-514."""
    dz, banked = _exner(orc, np.full((4, 5), 3.7e-3), np.full((5, 4), -1.1e-3))
    assert (dz == 0.0).all() and (banked == 0.0).all()


def test_exner_matches_the_hand_computed_divergence(orc):
    """test_sediment.py:517-526: -gain / dx / (1 - p)."""
    gain, dx, p = 2.0e-3, 50.0, 0.4
    qs_x = gain * np.arange(6, dtype=np.float64)[None, :] * np.ones((3, 1))
    dz, _ = _exner(orc, qs_x, np.zeros((4, 5)), dx=dx, porosity=p)
    assert np.allclose(dz, -gain / dx / (1.0 - p), rtol=1e-6)


def test_a_bound_banks_exactly_what_it_refuses(orc):
    """test_sediment.py:528-548."""
    gain, dx, p = 2.0e-3, 50.0, 0.4
    qs_x = gain * np.arange(4, dtype=np.float64)[None, :] * np.ones((1, 1))
    want = -gain / dx / (1.0 - p)
    frozen = np.zeros((1, 3), np.float32)
    dz, banked = _exner(orc, qs_x, np.zeros((2, 3)), dx=dx, porosity=p, lo=frozen, hi=frozen)
    assert (dz == 0.0).all()
    assert np.allclose(banked, want, rtol=1e-6)
    floor = np.full((1, 3), 0.5 * want, np.float32)
    dz, banked = _exner(orc, qs_x, np.zeros((2, 3)), dx=dx, porosity=p, lo=floor)
    assert np.allclose(dz, floor, rtol=1e-6)
    assert np.allclose(banked, want - floor, rtol=1e-5)


def test_the_bed_is_rebuilt_from_the_pristine_bed(orc):
    """test_sediment.py:551-580: a 1e-6 m increment on a 175 m bed survives."""
    z0 = np.full((2, 2), 175.0, np.float32)
    z = z0.copy()
    dz = np.full((2, 2), 1.0e-6 * 100, np.float64)
    orc.rebuild_bed(z0, dz, z)
    assert np.abs((z - z0) - 1e-4).max() <= np.spacing(np.float32(175.0))


def test_transport_integral_in_step_leaves_water_unchanged(orc):
    """morphology: the in-step hook reads faces and writes only the integral
    (processes/test_morphology.py:90-141): water is bitwise the unarmed run's."""
    rng = np.random.default_rng(0)
    z = (np.linspace(2, 0, 30)[None, :] * np.ones((6, 1))).astype(np.float32)
    a = orc.State(z, 5.0, depth=0.4, manning=0.03)
    b = orc.State(z, 5.0, depth=0.4, manning=0.03)
    b.arm_sediment(rng.uniform(0.001, 0.004, (6, 30)).astype(np.float32))
    for _ in range(40):
        dt = orc.li_compute_dt(a, 0.7, 2.0)
        orc.li_step(a, dt)
        orc.li_step(b, dt)
    assert np.array_equal(a.h, b.h) and np.array_equal(a.qx, b.qx)
    assert np.abs(b.qs_int_x).max() > 0  # it did transport
    assert np.all(b.qs_int_x[:, 0] == 0) and np.all(b.qs_int_x[:, -1] == 0)  # closed edges


def test_transport_integral_x_y_symmetry(orc):
    """test_sediment.py:328-371: a transposed reach transports the transposed integral."""
    z = (np.linspace(1, 0, 24)[None, :] * np.ones((5, 1))).astype(np.float32)
    a = orc.State(z, 4.0, depth=0.3, manning=0.03)
    b = orc.State(z.T.copy(), 4.0, depth=0.3, manning=0.03)
    a.arm_sediment(0.002)
    b.arm_sediment(0.002)
    for _ in range(20):
        orc.li_step(a, 0.5)
        orc.li_step(b, 0.5)
    np.testing.assert_allclose(a.qs_int_x, b.qs_int_y.T, rtol=1e-5, atol=1e-12)


@pytest.mark.parametrize("dx,porosity", [(5.0, 0.4), (0.3, 0.35)])
def test_exner_takes_float32_parameters(orc, dx, porosity):
    """sediment.py:384-392: dx and 1/(1-p) are wp.float32 kernel parameters widened
    inside, so a unit divergence moves the bed by exactly
    -f64(f32(1/(1-p))) / f64(f32(dx)), not by the float64 quotient."""
    qx = np.zeros((1, 2), np.float32)
    qx[0, 1] = 1.0
    qy = np.zeros((2, 1), np.float32)
    lo = np.full((1, 1), -np.inf, np.float32)
    hi = np.full((1, 1), np.inf, np.float32)
    dz, un = np.zeros((1, 1)), np.zeros((1, 1))
    orc.exner_update(qx, qy, lo, hi, dx, porosity, dz, un)
    want = -float(np.float32(1.0 / (1.0 - porosity))) / float(np.float32(dx))
    assert dz[0, 0] == want
    assert dz[0, 0] != -(1.0 / (1.0 - porosity)) / dx
