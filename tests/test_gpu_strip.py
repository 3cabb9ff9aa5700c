"""The grounded strip (strip.py:100-327) on the GPU against the oracle's restatement, and the
reference's own strip tests (test_strip.py) run on the GPU class."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def relative(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))


def cloud(n=1500, seed=0):
    rng = np.random.default_rng(seed)
    pos = np.stack((rng.uniform(0.05, 0.95, n), rng.uniform(-2, 2, n)), axis=0)
    return pos, rng.normal(0.2, 1.0, n)


@pytest.mark.parametrize("method", ["fmm", "direct"])
@pytest.mark.parametrize("softening", [0.0, 0.02])
def test_field_at_charges_matches_oracle(td3, oracle, method, softening):
    pos, q = cloud(3000, seed=1)
    # a few close pairs, so the softening has pairs to correct
    pos[:, 1::50] = pos[:, 0:-1:50][:, : pos[:, 1::50].shape[1]] + 0.003
    print(f"strip field_at_charges {method} eps={softening}", flush=True)
    strip = td3.GroundedStripField(1.0, k_e=1.0, method=method, softening=softening)
    phi, field = strip.field_at_charges(pos, q)
    rphi, rfield = oracle.strip_field_at_charges(1.0, pos, q, k_e=1.0, method=method,
                                                 softening=softening)
    assert relative(phi, rphi) < 1e-12 and relative(field, rfield) < 1e-12
    _, field_only = strip.field_at_charges(pos, q, potential=False)
    assert relative(field_only, rfield) < 1e-12


@pytest.mark.parametrize("method", ["fmm", "direct"])
def test_probe_matches_oracle(td3, oracle, method):
    pos, q = cloud(2000, seed=2)
    rng = np.random.default_rng(3)
    probes = np.stack((rng.uniform(0.02, 0.98, 500), rng.uniform(-2.5, 2.5, 500)))
    probes[:, :40] = pos[:, :40] + 0.004  # probes within the softening cutoff of charges
    print(f"strip probe {method}", flush=True)
    for softening in (0.0, 0.01):
        strip = td3.GroundedStripField(1.0, k_e=1.0, method=method, softening=softening)
        phi, field = strip.probe(pos, q, probes)
        rphi, rfield = oracle.strip_probe(1.0, pos, q, probes, k_e=1.0, method=method,
                                          softening=softening)
        assert relative(phi, rphi) < 1e-12 and relative(field, rfield) < 1e-12


def test_closed_form_and_image_force(td3):
    """test_strip.py:38-60: grounded plates, the Green's function, a lone charge."""
    strip = td3.GroundedStripField(2.0, k_e=1.0, method="direct")
    source = np.array([[0.7], [0.3]])
    walls = strip.potential(source, [1.0], [[1e-12, 2.0 - 1e-12], [0.5, -1.0]])
    np.testing.assert_allclose(walls, 0.0, atol=1e-10)
    probe = np.array([1.4, -0.6])
    z, z0, a = complex(*probe), complex(0.7, 0.3), 2.0 / np.pi
    expected = -2 * np.log(abs(np.sin((z - z0) / (2 * a)) / np.sin((z + np.conj(z0)) / (2 * a))))
    assert strip.potential(source, [1.0], probe[:, None])[0] == pytest.approx(expected)
    lone = td3.GroundedStripField(1.0, k_e=1.0)
    _, middle = lone.field_at_charges([[0.5], [0.0]], [1.0])
    np.testing.assert_allclose(middle[:, 0], 0.0, atol=1e-10)
    _, near_left = lone.field_at_charges([[0.2], [0.0]], [1.0])
    assert near_left[0, 0] == pytest.approx(-np.pi / np.tan(np.pi * 0.2), rel=1e-10)


def test_softening_is_local_and_no_net_y_force(td3):
    """test_strip.py:86-126."""
    far = np.array([[0.3, 0.6], [0.0, 0.4]])
    q = np.array([1.0, -2.0])
    bare = td3.GroundedStripField(1.0, k_e=1.0, method="direct")
    soft = td3.GroundedStripField(1.0, k_e=1.0, method="direct", softening=0.01)
    for x, y in zip(bare.field_at_charges(far, q), soft.field_at_charges(far, q)):
        np.testing.assert_array_equal(x, y)
    close = np.array([[0.3, 0.305], [0.0, 0.002]])
    _, e_bare = bare.field_at_charges(close, q)
    _, e_soft = soft.field_at_charges(close, q)
    d = close[:, 0] - close[:, 1]
    r2 = d @ d
    change = 2 * (1 / (r2 + 1e-4) - 1 / r2) * d
    np.testing.assert_allclose(e_soft[:, 0], e_bare[:, 0] + q[1] * change)
    np.testing.assert_allclose(e_soft[:, 1], e_bare[:, 1] - q[0] * change)
    pos, qq = cloud(800, seed=4)
    for method in ("direct", "fmm"):
        _, field = td3.GroundedStripField(1.0, k_e=1.0, method=method,
                                          softening=0.02).field_at_charges(pos, qq)
        force = qq * field
        assert abs(force[1].sum()) < 1e-6 * np.abs(force[1]).sum()


def test_validation_and_device_inputs(td3):
    strip = td3.GroundedStripField(1.0)
    with pytest.raises(ValueError, match="between the plates"):
        strip.field_at_charges([[1.0], [0.0]], [1.0])
    with pytest.raises(ValueError, match="method"):
        td3.GroundedStripField(1.0, method="tree")
    with pytest.raises(ValueError, match="softening"):
        td3.GroundedStripField(1.0, softening=-1)
    phi, field = strip.field_at_charges(np.zeros((2, 0)), [])
    assert phi.shape == (0,) and field.shape == (2, 0)
    with pytest.raises(ValueError, match="Coincident"):
        td3.GroundedStripField(1.0, softening=0.01, method="direct").field_at_charges(
            [[0.3, 0.3], [0.1, 0.1]], [1.0, 1.0])
    with pytest.raises(ValueError, match="oincide"):
        td3.GroundedStripField(1.0, softening=0.01).probe([[0.3], [0.1]], [1.0], [[0.3], [0.1]])
    pos, q = cloud(500, seed=7)
    t_pos, t_q = torch.tensor(pos, device="cuda"), torch.tensor(q, device="cuda")
    phi, field = td3.GroundedStripField(1.0, k_e=1.0, softening=0.01).field_at_charges(t_pos, t_q)
    assert phi.is_cuda and field.is_cuda
