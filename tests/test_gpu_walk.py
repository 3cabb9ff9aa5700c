"""GPU Barnes-Hut walk (FlatQuadtree.evaluate / field_at) against the oracle and direct sums."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relative(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def config1(seed=1, n=5000):
    """The Barnes-Hut workload: uniform [-1, 1]^2, charges +-1, Coulomb, theta 0.5, eps 1e-3."""
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, (2, n)), rng.choice([-1.0, 1.0], n)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config1_walk(td3, oracle, precision):
    pos, q = config1()
    n = pos.shape[1]
    own = np.arange(n)
    print(f"walk config1 precision={precision}", flush=True)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), precision=precision)
    phi, field = tree.evaluate(pos, exclude=own, theta=0.5, softening=1e-3)
    ref = oracle.OracleTree(pos, q, (0, 0, 1))
    rphi, rfield = ref.evaluate(pos, exclude=own, theta=0.5, softening=1e-3)
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert relative(phi, rphi) < tol and relative(field, rfield) < tol
    dphi, dfield = oracle.direct_field(pos, pos, q, exclude=own, softening=1e-3)
    # the walk's truncation error against the direct sum (quadrupole, theta 0.5)
    assert relative(field, dfield) < 2e-2 and relative(phi, dphi) < 1e-2


@pytest.mark.parametrize("theta", [0.0, 0.5, 1.2])
@pytest.mark.parametrize("kernel", ["coulomb", "log"])
def test_walk_matches_oracle(td3, oracle, theta, kernel):
    rng = np.random.default_rng(4)
    pos = np.ascontiguousarray(rng.normal(0, 0.3, (300, 2)).T).clip(-0.99, 0.99)
    pos[:, :30] = pos[:, :1]  # a coincident clump exercises max_depth buckets
    q = np.where(rng.random(300) < 0.5, -1.0, 1.0) * rng.uniform(0.5, 2, 300)
    q[rng.random(300) < 0.1] = 0.0
    own = np.arange(300)
    for leaf, depth in ((1, 10), (8, 48)):
        tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), leaf_capacity=leaf, max_depth=depth,
                                kernel=kernel)
        ref = oracle.OracleTree(pos, q, (0, 0, 1), leaf_capacity=leaf, max_depth=depth,
                                kernel=kernel)
        got = tree.evaluate(pos, exclude=own, theta=theta, softening=0.01)
        want = ref.evaluate(pos, exclude=own, theta=theta, softening=0.01)
        scale = np.abs(q).sum() / 0.01
        np.testing.assert_allclose(got[0], want[0], rtol=0, atol=1e-12 * scale)
        np.testing.assert_allclose(got[1], want[1], rtol=0, atol=1e-12 * scale / 0.01)
        if theta == 0.0:
            direct = oracle.direct_field(pos, pos, q, exclude=own, softening=0.01, kernel=kernel)
            np.testing.assert_allclose(got[1], direct[1], rtol=0, atol=1e-12 * scale / 0.01)


@pytest.mark.parametrize("order,tolerance", [(4, 1e-3), (8, 1e-5), (12, 2e-7)])
def test_log_walk_converges(td3, oracle, order, tolerance):
    # test_log_tree.py:19-28
    rng = np.random.default_rng(0)
    pos, q = rng.normal(0, 0.2, (2, 3000)).clip(-0.99, 0.99), rng.normal(0.3, 1, 3000)
    probes = np.random.default_rng(1).uniform(-1, 1, (2, 200))
    phi, field = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", order=order).evaluate(
        probes, theta=0.6)
    ref_phi, ref_field = oracle.direct_field(probes, pos, q, kernel="log")
    assert relative(phi, ref_phi) < tolerance
    assert relative(field, ref_field) < tolerance


def test_walk_edges(td3):
    pos = np.array([[0.1, 0.5], [0.2, -0.3]])
    tree = td3.FlatQuadtree(pos, np.ones(2), td3.Quad(0, 0, 1))
    with pytest.raises(ValueError, match="Coincident"):
        tree.evaluate(pos[:, :1])
    phi, field = tree.evaluate(pos[:, :1], exclude=[0])
    assert np.isfinite(phi[0])
    nophi, same = tree.evaluate(pos, exclude=[0, 1], potential=False)
    assert not np.any(nophi)
    phi, field = tree.evaluate(np.zeros((2, 0)))
    assert field.shape == (2, 0)
    # test_log_tree.py:31-38: one line charge, field_at with k_e and a reference radius
    line = td3.FlatQuadtree([[0.0], [0.0]], [2.0], td3.Quad(0, 0, 1), kernel="log")
    phi, field = line.field_at([[3.0], [4.0]], k_e=1.0, reference_radius=2.0)
    assert phi[0] == pytest.approx(-2 * 2.0 * np.log(5.0 / 2.0))
    np.testing.assert_allclose(field[:, 0], 2 * 2.0 * np.array([3.0, 4.0]) / 25)


def test_walk_torch_inputs_stay_on_device(td3):
    import torch

    pos, q = config1(n=2000)
    tpos = torch.from_numpy(pos).cuda()
    tree = td3.FlatQuadtree(tpos, torch.from_numpy(q).cuda(), td3.Quad(0, 0, 1))
    phi, field = tree.evaluate(tpos, exclude=torch.arange(2000, device="cuda"), softening=1e-3)
    assert phi.is_cuda and field.is_cuda and field.shape == (2, 2000)


def test_mixed_log_walk_deep_tree_high_order(td3):
    """Cells 2^-40 wide at order 40: the mixed walk's radius-scaled moments a_k h^-k stay
    finite (h^-k alone overflows FP64 there), so the fields stay finite and within FP32
    rounding of the FP64 walk's."""
    rng = np.random.default_rng(9)
    tight = 0.3 + 1e-12 * rng.uniform(-1, 1, (2, 64))
    pos = np.concatenate((tight, rng.uniform(-1, 1, (2, 500))), axis=1)
    q = rng.choice([-1.0, 1.0], pos.shape[1])
    targets = rng.uniform(-1, 1, (2, 300))
    print("deep tree order 40 mixed walk", flush=True)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", order=40, leaf_capacity=1)
    assert int(np.max(tree.node_depth)) >= 38
    phi, field = tree.evaluate(targets, theta=0.5, precision="mixed")
    rphi, rfield = tree.evaluate(targets, theta=0.5, precision="fp64")
    assert np.all(np.isfinite(phi)) and np.all(np.isfinite(field))
    assert relative(phi, rphi) < 1e-5 and relative(field, rfield) < 1e-5
