"""Pins the oracle's grounded strip (strip_field_at_charges, strip_probe, soften_pairs,
restating strip.py:100-327 and kernels/soften_pairs) by what fixes it independently of its
own formulas: the strip's closed-form Green's function and image force (strip.py:56-97,
the reference's test_strip.py), grounded plates, sheets of line charges against the 1D
solution, zero net force along y, and the softening as a local pair correction."""

import numpy as np
import pytest


def cloud(n=1500, seed=0):
    """test_strip.py:22-25."""
    rng = np.random.default_rng(seed)
    pos = np.stack((rng.uniform(0.05, 0.95, n), rng.uniform(-2, 2, n)), axis=0)
    return pos, rng.normal(0.2, 1.0, n)


def sheet_potential(sheets, sigma, points, length, epsilon):
    """Charged sheets between grounded walls at 0 and length (1D), test_strip.py:12-19."""
    field_0 = -np.sum(sigma * (length - sheets)) / (epsilon * length)
    below = sheets[None, :] < points[:, None]
    return (-field_0 * points
            - np.sum(below * sigma * (points[:, None] - sheets), axis=1) / epsilon)


def test_green_function_and_grounded_plates(oracle):
    rng = np.random.default_rng(1)
    width = 2.3
    a = width / np.pi
    for _ in range(6):
        x0, y0, x, y = rng.uniform(0.1, 2.2), rng.uniform(-1, 1), rng.uniform(0.1, 2.2), \
            rng.uniform(-1, 1)
        q = rng.normal()
        want = -q * np.log((np.sin((x - x0) / (2 * a)) ** 2 + np.sinh((y - y0) / (2 * a)) ** 2)
                           / (np.sin((x + x0) / (2 * a)) ** 2 + np.sinh((y - y0) / (2 * a)) ** 2))
        for method in ("direct", "fmm"):
            phi, _ = oracle.strip_probe(width, [[x0], [y0]], [q], [[x], [y]], k_e=1.0,
                                        method=method)
            assert phi[0] == pytest.approx(want, rel=1e-10, abs=1e-12)
    walls, _ = oracle.strip_probe(2.0, [[0.7], [0.3]], [1.0], [[1e-12, 2.0 - 1e-12], [0.5, -1.0]],
                                  k_e=1.0, method="direct")
    np.testing.assert_allclose(walls, 0.0, atol=1e-10)


def test_image_force_of_a_lone_charge(oracle):
    # strip.py:83-97: E_x = -(pi k_e q / L) cot(pi x / L), E_y = 0
    for width, x, q in ((1.0, 0.2, 1.0), (2.5, 0.4, -0.7), (2.0, 1.0, 3.0)):
        for method in ("direct", "fmm"):
            _, field = oracle.strip_field_at_charges(width, [[x], [0.3]], [q], k_e=1.0,
                                                     method=method)
            want = -np.pi * q / width / np.tan(np.pi * x / width)
            assert field[0, 0] == pytest.approx(want, rel=1e-10, abs=1e-12)
            assert abs(field[1, 0]) < 1e-12


def test_fmm_matches_direct_and_no_net_y_force(oracle):
    pos, q = cloud(1200)
    phi_f, field_f = oracle.strip_field_at_charges(1.0, pos, q, k_e=1.0)
    phi_d, field_d = oracle.strip_field_at_charges(1.0, pos, q, k_e=1.0, method="direct")
    assert np.abs(phi_f - phi_d).max() < 1e-6 * np.abs(phi_d).max()
    assert np.abs(field_f - field_d).max() < 1e-5 * np.abs(field_d).max()
    force = q * field_d
    assert abs(force[1].sum()) < 1e-6 * np.abs(force[1]).sum()


def test_rows_of_line_charges_reproduce_sheets(oracle):
    rows = np.linspace(-20, 20, 800, endpoint=False)
    sheets, sigma = np.array([0.3, 0.6]), np.array([1.0, -0.4])
    lines = np.array([[x, y] for x in sheets for y in rows]).T
    per_line = np.repeat(sigma * 40 / len(rows), len(rows))
    xs = np.linspace(0.02, 0.98, 9)
    probes = np.stack((xs, np.zeros_like(xs)), axis=0)
    two_d, _ = oracle.strip_probe(1.0, lines, per_line, probes, k_e=1.0)
    one_d = sheet_potential(sheets, sigma, xs, 1.0, 1 / (4 * np.pi))
    assert np.abs(two_d - one_d).max() < 5e-3 * np.abs(one_d).max()


def test_softening_is_a_local_pair_correction(oracle):
    far = np.array([[0.3, 0.6], [0.0, 0.4]])
    q = np.array([1.0, -2.0])
    for a, b in zip(oracle.strip_field_at_charges(1.0, far, q, k_e=1.0, method="direct"),
                    oracle.strip_field_at_charges(1.0, far, q, k_e=1.0, method="direct",
                                                  softening=0.01)):
        np.testing.assert_array_equal(a, b)
    close = np.array([[0.3, 0.305], [0.0, 0.002]])
    _, e_bare = oracle.strip_field_at_charges(1.0, close, q, k_e=1.0, method="direct")
    p_soft, e_soft = oracle.strip_field_at_charges(1.0, close, q, k_e=1.0, method="direct",
                                                   softening=0.01)
    p_bare, _ = oracle.strip_field_at_charges(1.0, close, q, k_e=1.0, method="direct")
    d = close[:, 0] - close[:, 1]
    r2 = d @ d
    change = 2 * (1 / (r2 + 1e-4) - 1 / r2) * d
    np.testing.assert_allclose(e_soft[:, 0], e_bare[:, 0] + q[1] * change, rtol=1e-13)
    np.testing.assert_allclose(e_soft[:, 1], e_bare[:, 1] - q[0] * change, rtol=1e-13)
    dphi = -np.log((r2 + 1e-4) / r2)
    np.testing.assert_allclose(p_soft, p_bare + dphi * q[::-1], rtol=1e-13)
    # probes: the same difference, source by source
    probe = np.array([[0.3021], [0.0013]])
    pb, fb = oracle.strip_probe(1.0, close, q, probe, k_e=1.0, method="direct")
    ps, fs = oracle.strip_probe(1.0, close, q, probe, k_e=1.0, method="direct", softening=0.01)
    dd = probe - close
    rr = (dd ** 2).sum(0)
    np.testing.assert_allclose(ps, pb - np.sum(q * np.log((rr + 1e-4) / rr)), rtol=1e-12)
    np.testing.assert_allclose(fs[:, 0], fb[:, 0] + np.sum(q * (2 / (rr + 1e-4) - 2 / rr) * dd,
                                                           axis=1), rtol=1e-12)


def test_validation(oracle):
    with pytest.raises(ValueError, match="between the plates"):
        oracle.strip_field_at_charges(1.0, [[1.0], [0.0]], [1.0])
    phi, field = oracle.strip_field_at_charges(1.0, np.zeros((2, 0)), [])
    assert phi.shape == (0,) and field.shape == (2, 0)
    with pytest.raises(ValueError, match="Coincident"):
        oracle.strip_field_at_charges(1.0, [[0.3, 0.3], [0.1, 0.1]], [1.0, 1.0], softening=0.01)
