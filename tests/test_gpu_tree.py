"""GPU tree build and moments against the CPU oracle.

The tree must be bit-identical (test_tree_build.py:38-53, test_cuda_kernels.py:61-78 of the
reference require array_equal); moments agree to rounding (1e-12 after scaling).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TREE_ARRAYS = ("perm", "node_start", "node_end", "node_child", "node_depth", "node_center",
               "node_half", "levels")


def clouds():
    """test_tree_build.py:22-33 clouds plus the benchmark's uniform+cluster mix."""
    rng = np.random.default_rng(5)
    angle = rng.uniform(0, 2 * np.pi, 3000)
    mix = np.concatenate(
        (rng.uniform(-1, 1, (2, 20000)),
         (rng.uniform(-0.8, 0.8, (2, 16))[:, :, None] + rng.normal(0, 0.02, (2, 16, 1250))
          ).reshape(2, -1)), axis=1).clip(-1, 1)
    return {
        "uniform": rng.uniform(-0.99, 0.99, (2, 4000)),
        "circle": 0.9 * np.stack((np.cos(angle), np.sin(angle))),
        "clustered": np.clip(rng.normal(0.3, 0.01, (2, 2000)), -1, 1),
        "lattice": np.stack(np.meshgrid(np.linspace(-0.5, 0.5, 33),
                                        np.linspace(-0.5, 0.5, 33))).reshape(2, -1),
        "coincident": np.full((2, 40), 0.25),
        "mix": mix,
        "edges": np.array([[-1.0, 1.0, 0.0, 0.0, 0.5, -0.5, 1.0, -1.0],
                           [1.0, -1.0, 0.0, 0.5, 0.0, -0.5, 1.0, -1.0]]),
    }


def assert_same_tree(tree, ref):
    assert tree.node_count == ref.node_count
    for name in TREE_ARRAYS:
        a = np.asarray(getattr(tree, name))
        b = np.asarray(getattr(ref, name))
        assert a.shape == b.shape, name
        np.testing.assert_array_equal(a, b, err_msg=name)


@pytest.mark.parametrize("name", list(clouds()))
@pytest.mark.parametrize("leaf", [1, 6, 8, 40])
def test_tree_bit_identical(td3, oracle, name, leaf):
    pos = clouds()[name]
    q = np.ones(pos.shape[1])
    print(f"tree cloud={name} n={pos.shape[1]} leaf={leaf}", flush=True)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), leaf_capacity=leaf)
    ref = oracle.OracleTree(pos, q, (0, 0, 1), leaf_capacity=leaf)
    assert_same_tree(tree, ref)


@pytest.mark.parametrize("max_depth", [1, 5, 31, 32, 33, 48, 64])
def test_tree_depth_limits(td3, oracle, max_depth):
    pos = np.array([[0.3, 0.3, -0.5, 0.3000000001], [0.3, 0.3, 0.2, 0.3]])
    q = np.array([1.0, 1.0, -1.0, 2.0])
    print(f"tree max_depth={max_depth}", flush=True)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), leaf_capacity=1, max_depth=max_depth)
    ref = oracle.OracleTree(pos, q, (0, 0, 1), leaf_capacity=1, max_depth=max_depth)
    assert_same_tree(tree, ref)


def test_coincident_chain(td3):
    # test_flat_tree.py:92-102
    tree = td3.FlatQuadtree([[0.3, 0.3, -0.5], [0.3, 0.3, 0.2]], [1, 1, -1], td3.Quad(0, 0, 1),
                            leaf_capacity=1, max_depth=48)
    assert tree.node_count == 1 + 4 * 48
    assert tree.node_depth.max() == 48
    assert tree.root.count == 3


def test_non_dyadic_root(td3, oracle):
    # the target tree of fmm_at: half * (1 + 1e-12) + 1e-300 (flat_tree.py:1135-1137)
    rng = np.random.default_rng(9)
    targets = rng.uniform(-0.7, 0.9, (2, 5000))
    lo, hi = targets.min(axis=1), targets.max(axis=1)
    centre = 0.5 * (lo + hi)
    half = 0.5 * float(np.max(hi - lo)) * (1 + 1e-12) + 1e-300
    tree = td3.FlatQuadtree(targets, np.zeros(5000), td3.Quad(centre[0], centre[1], half),
                            leaf_capacity=8)
    ref = oracle.OracleTree(targets, np.zeros(5000), (centre[0], centre[1], half),
                            leaf_capacity=8)
    assert_same_tree(tree, ref)


def test_large_mix_tree(td3, oracle):
    rng = np.random.default_rng(11)
    n = 200000
    pos = np.concatenate(
        (rng.uniform(-1, 1, (2, n // 2)),
         (rng.uniform(-0.8, 0.8, (2, 16))[:, :, None] + rng.normal(0, 0.02, (2, 16, n // 32))
          ).reshape(2, -1)), axis=1).clip(-1, 1)
    q = rng.choice([-1.0, 1.0], n)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log")
    ref = oracle.OracleTree(pos, q, (0, 0, 1), kernel="log")
    assert_same_tree(tree, ref)


@pytest.mark.parametrize("kernel", ["coulomb", "log"])
def test_moments_match_oracle(td3, oracle, kernel):
    pos = clouds()["mix"]
    rng = np.random.default_rng(3)
    q = rng.normal(0.3, 1.0, pos.shape[1])
    q[::9] = 0.0
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel=kernel, order=12)
    ref = oracle.OracleTree(pos, q, (0, 0, 1), kernel=kernel, order=12)
    scale = np.abs(q).sum()
    np.testing.assert_allclose(tree.node_qabs, ref.node_qabs, rtol=1e-12, atol=1e-12 * scale)
    np.testing.assert_allclose(tree.node_q, ref.node_q, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(tree.node_p, ref.node_p, rtol=0, atol=1e-12 * scale)
    if kernel == "log":
        for k in range(13):
            np.testing.assert_allclose(tree.moment_re[:, k], ref.moment_re[:, k], rtol=0,
                                       atol=1e-12 * scale * 2.0 ** k)
            np.testing.assert_allclose(tree.moment_im[:, k], ref.moment_im[:, k], rtol=0,
                                       atol=1e-12 * scale * 2.0 ** k)
    else:
        np.testing.assert_allclose(tree.node_s, ref.node_s, rtol=0, atol=1e-12 * scale)
    other = rng.normal(size=pos.shape[1])
    tree.set_charges(other)
    fresh = td3.FlatQuadtree(pos, other, td3.Quad(0, 0, 1), kernel=kernel, order=12)
    np.testing.assert_array_equal(tree.node_qabs, fresh.node_qabs)
    np.testing.assert_array_equal(tree.moment_re, fresh.moment_re)


def test_validation_and_empty(td3):
    with pytest.raises(ValueError, match=r"\(2, N\)"):
        td3.FlatQuadtree(np.zeros((3, 2)), np.zeros(3), td3.Quad(0, 0, 1))
    with pytest.raises(ValueError, match="^Particle lies outside tree bounds$"):
        td3.FlatQuadtree([[2.0], [0.0]], [1.0], td3.Quad(0, 0, 1))
    # flat_tree.py:785-794: finite first, then the bounds; the closed square is inside
    many = np.random.default_rng(0).uniform(-1, 1, (2, 5000))
    for bad, message in ((np.nan, "finite"), (np.inf, "finite"), (1.0 + 1e-15, "outside"),
                         (-1.5, "outside")):
        pos = many.copy()
        pos[1, 4321] = bad
        with pytest.raises(ValueError, match=message):
            td3.FlatQuadtree(pos, np.ones(5000), td3.Quad(0, 0, 1))
    pos = many.copy()
    pos[:, :4] = [[-1.0, 1.0, -1.0, 1.0], [-1.0, -1.0, 1.0, 1.0]]  # the corners
    assert td3.FlatQuadtree(pos, np.ones(5000), td3.Quad(0, 0, 1)).node_count > 1
    pos[0, 7], pos[1, 9] = np.nan, 3.0
    with pytest.raises(ValueError, match="must be finite"):
        td3.FlatQuadtree(pos, np.ones(5000), td3.Quad(0, 0, 1))
    with pytest.raises(ValueError, match="kernel"):
        td3.FlatQuadtree([[0.0], [0.0]], [1.0], td3.Quad(0, 0, 1), kernel="yukawa")
    with pytest.raises(ValueError, match="order"):
        td3.FlatQuadtree([[0.0], [0.0]], [1.0], td3.Quad(0, 0, 1), order=65)
    empty = td3.FlatQuadtree(np.zeros((2, 0)), np.zeros(0), td3.Quad(0, 0, 1))
    assert empty.node_count == 1
    empty.set_charges(np.zeros(0))
