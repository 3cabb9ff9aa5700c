"""Pins the oracle's particle push (simulation.py:812-835) by the exact solution of the
gyration, the B_z -> 0 limit (velocity Verlet) and its invariants."""

import numpy as np
import pytest
import sympy

from oracle import oracle


def state(n, seed):
    rng = np.random.default_rng(seed)
    return dict(positions=rng.uniform(-1, 1, (2, n)), velocities=rng.normal(0, 1, (2, n)),
                forces=np.zeros((2, n)), electric=rng.normal(0, 1, (2, n)),
                masses=rng.uniform(0.5, 2, n), charges=rng.normal(0, 1, n),
                fixed=rng.random(n) < 0.2)


def test_drift_is_the_exact_gyration():
    # dv/dt = w (v_y, -v_x) with w = q B_z / m: v(t) = R(-w t) v0, r(t) = r0 + int v
    t, w, vx, vy = sympy.symbols("t w vx vy", real=True)
    vel = (vx * sympy.cos(w * t) + vy * sympy.sin(w * t), vy * sympy.cos(w * t) - vx * sympy.sin(w * t))
    assert sympy.simplify(sympy.diff(vel[0], t) - w * vel[1]) == 0
    assert sympy.simplify(sympy.diff(vel[1], t) + w * vel[0]) == 0
    tau = sympy.symbols("tau", positive=True)
    drift = [sympy.integrate(v, (t, 0, tau)) for v in vel]
    s = state(40, 1)
    s["electric"][:] = 0
    s["fixed"][:] = False
    p0, v0 = s["positions"].copy(), s["velocities"].copy()
    dt, b_z = 0.37, 1.9
    oracle.push(**s, dt=dt, b_z=b_z, first_half=True)
    for i in range(40):
        subs = {w: s["charges"][i] / s["masses"][i] * b_z, vx: v0[0, i], vy: v0[1, i], tau: dt}
        # 40 digits: (1 - cos(w tau)) / w cancels for small w
        want_r = [float(d.evalf(40, subs=subs)) for d in drift]
        want_v = [float(v.evalf(40, subs=subs | {t: dt})) for v in vel]
        np.testing.assert_allclose(s["positions"][:, i] - p0[:, i], want_r, rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(s["velocities"][:, i], want_v, rtol=1e-12, atol=1e-14)


def test_zero_field_limit_is_velocity_verlet():
    s = state(50, 2)
    p0, v0, e = s["positions"].copy(), s["velocities"].copy(), s["electric"].copy()
    dt = 0.01
    oracle.push(**s, dt=dt, b_z=0.0, first_half=True)
    mobile = ~s["fixed"]
    half = v0 + 0.5 * dt * e / s["masses"]
    np.testing.assert_allclose(s["velocities"][:, mobile], half[:, mobile], rtol=1e-15)
    np.testing.assert_allclose(s["positions"][:, mobile], p0[:, mobile] + dt * half[:, mobile],
                               rtol=1e-14)
    np.testing.assert_array_equal(s["positions"][:, ~mobile], p0[:, ~mobile])
    np.testing.assert_array_equal(s["velocities"][:, ~mobile], v0[:, ~mobile])


def test_rotation_keeps_speed_and_tiny_angles_are_finite():
    s = state(30, 3)
    s["electric"][:] = 0
    speed = np.hypot(*s["velocities"])
    oracle.push(**s, dt=1e-3, b_z=1e-9, first_half=True)
    assert np.all(np.isfinite(s["positions"]))
    np.testing.assert_allclose(np.hypot(*s["velocities"]), speed, rtol=1e-15)
    oracle.push(**s, dt=0.5, b_z=40.0, first_half=True)
    np.testing.assert_allclose(np.hypot(*s["velocities"]), speed, rtol=1e-14)


def test_closing_half_sets_lorentz_forces():
    s = state(20, 4)
    v0 = s["velocities"].copy()
    dt, b_z = 0.2, 0.8
    oracle.push(**s, dt=dt, b_z=b_z, first_half=False)
    v = np.where(s["fixed"], v0, v0 + 0.5 * dt * s["electric"] / s["masses"])
    np.testing.assert_allclose(s["velocities"], v, rtol=1e-15)
    want = s["electric"] + s["charges"] * b_z * np.stack((v[1], -v[0]))
    np.testing.assert_allclose(s["forces"], want, rtol=1e-15)
    # a planted sign error in the magnetic term is caught
    wrong = s["electric"] + s["charges"] * b_z * np.stack((-v[1], v[0]))
    assert not np.allclose(s["forces"], wrong)


@pytest.mark.parametrize("first_half", [True, False])
def test_uncharged_particle_in_field_free_space_moves_straight(first_half):
    s = state(5, 5)
    s["charges"][:] = 0
    s["electric"][:] = 0
    s["fixed"][:] = False
    p0, v0 = s["positions"].copy(), s["velocities"].copy()
    oracle.push(**s, dt=0.1, b_z=3.0, first_half=first_half)
    np.testing.assert_allclose(s["velocities"], v0, rtol=1e-15)
    if first_half:
        np.testing.assert_allclose(s["positions"], p0 + 0.1 * v0, rtol=1e-15)
