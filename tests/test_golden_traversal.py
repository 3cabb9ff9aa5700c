"""The reference's own tree arrays and interaction lists (tests/golden/traversal.npz, written
by tests/golden/make_traversal_vectors.py from the reference's quadtree_build,
fmm_traverse and dual_traverse kernels) against the oracle (CPU) and the GPU.

The traversal's pair sets are what no mathematical check fixes (any cover of the particle
pairs by separated cells would pass the coverage tests), so these vectors pin them: a
changed split rule or acceptance comparison changes the sets."""

import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
TREE = ("node_center", "node_half", "node_start", "node_end", "node_child", "node_depth", "perm",
        "levels")


@pytest.fixture(scope="module")
def golden():
    with np.load(os.path.join(HERE, "golden", "traversal.npz")) as data:
        cases = {}
        for key in data.files:
            case, name = key.split("/")
            cases.setdefault(case, {})[name] = data[key]
    return cases


def mutual_cases(golden):
    return [name for name in golden if not name.startswith("dual")]


def sorted_pairs(pairs):
    pairs = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    return pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]


def as_directed(pairs, self_once):
    """Both directions of every mutual pair (a self pair once), sorted."""
    pairs = np.asarray(pairs, dtype=np.int64)
    back = pairs[:, ::-1]
    if self_once:
        back = back[pairs[:, 0] != pairs[:, 1]]
    return sorted_pairs(np.concatenate((pairs, back)))


def rows_as_pairs(rows, cols):
    """(row node, column node) of every entry of a row set, sorted."""
    row = np.asarray(rows.cpu()).astype(np.int64)
    col = np.asarray(cols.cpu()).astype(np.int64)[: row[-1]]
    dst = np.repeat(np.arange(len(row) - 1), np.diff(row))
    return sorted_pairs(np.stack((dst, col), axis=1))


def same_tree(tree, case, prefix=""):
    assert tree.node_count == len(case[f"{prefix}node_half"])
    for name in TREE:
        got = getattr(tree, name)
        got = got.cpu().numpy() if hasattr(got, "cpu") else np.asarray(got)
        want = case[f"{prefix}{name}"]
        assert got.shape == want.shape, name
        assert np.array_equal(got, want.astype(got.dtype)), name


def test_golden_vectors_are_whole(golden):
    assert len(mutual_cases(golden)) == 4 and {"dual_leaf8", "dual_leaf4"} <= set(golden)
    for case in golden.values():
        assert len(case["m2l"]) > 1000 and len(case["p2p"]) > 1000


def test_oracle_tree_equals_reference_build(golden, oracle):
    for name in mutual_cases(golden):
        case = golden[name]
        leaf, _, x, y, h = case["params"]
        n = case["pos"].shape[1]
        tree = oracle.OracleTree(case["pos"], np.ones(n), (x, y, h), leaf_capacity=int(leaf))
        same_tree(tree, case)


def test_oracle_target_tree_equals_reference_build(golden, oracle):
    # fmm_at's target tree: a non-dyadic root from the targets' bounds
    for name in ("dual_leaf8", "dual_leaf4"):
        case = golden[name]
        leaf, _, x, y, h = case["params"]
        m = case["targets"].shape[1]
        tree = oracle.OracleTree(case["targets"], np.zeros(m), (x, y, h), leaf_capacity=int(leaf),
                                 kernel="coulomb")
        same_tree(tree, case, "t_")


def test_oracle_traversal_equals_reference_lists(golden, oracle):
    for name in mutual_cases(golden):
        case = golden[name]
        leaf, theta, x, y, h = case["params"]
        n = case["pos"].shape[1]
        tree = oracle.OracleTree(case["pos"], np.ones(n), (x, y, h), kernel="log",
                                 leaf_capacity=int(leaf))
        m2l, p2p = tree.traverse(float(theta))
        # the same pairs, each in the reference's orientation
        assert np.array_equal(sorted_pairs(m2l), case["m2l"]), name
        assert np.array_equal(sorted_pairs(p2p), case["p2p"]), name


def test_oracle_dual_traversal_equals_reference_lists(golden, oracle):
    for name in ("dual_leaf8", "dual_leaf4"):
        case = golden[name]
        leaf, theta = case["params"][:2]
        n = case["sources"].shape[1]
        q = np.random.default_rng(0).choice([-1.0, 1.0], n)
        tree = oracle.OracleTree(case["sources"], q, (0.0, 0.0, 1.0), kernel="log")
        same_tree(tree, case, "s_")
        tree.fmm_at(case["targets"], theta=float(theta), leaf_capacity=int(leaf))
        m2l, p2p = tree.dual_lists
        assert np.array_equal(sorted_pairs(m2l), case["m2l"]), name
        assert np.array_equal(sorted_pairs(p2p), case["p2p"]), name


@pytest.mark.gpu
def test_gpu_tree_and_rows_equal_reference(golden, td3):
    for name in mutual_cases(golden):
        case = golden[name]
        leaf, theta, x, y, h = case["params"]
        n = case["pos"].shape[1]
        print(f"golden rows {name}", flush=True)
        tree = td3.FlatQuadtree(case["pos"], np.ones(n), td3.Quad(x, y, h), kernel="log",
                                leaf_capacity=int(leaf))
        same_tree(tree, case)
        (mrow, mcol, _), (prow, pcol, _), counts = tree._lists(float(theta))
        assert counts == (len(case["m2l"]), len(case["p2p"]))
        assert np.array_equal(rows_as_pairs(mrow, mcol), as_directed(case["m2l"], False))
        assert np.array_equal(rows_as_pairs(prow, pcol), as_directed(case["p2p"], True))
        tree.fmm(theta=float(theta))
        assert tree.fmm_pairs == counts


@pytest.mark.gpu
def test_gpu_dual_lists_equal_reference(golden, td3):
    import torch

    for name in ("dual_leaf8", "dual_leaf4"):
        case = golden[name]
        leaf, theta = case["params"][:2]
        n = case["sources"].shape[1]
        print(f"golden dual lists {name}", flush=True)
        tree = td3.FlatQuadtree(case["sources"], np.ones(n), td3.Quad(0, 0, 1), kernel="log")
        same_tree(tree, case, "s_")
        targets = torch.tensor(case["targets"], device="cuda")
        ttree, ((mrow, mcol, _), (prow, pcol, _), counts) = tree._target_plan(
            targets, float(theta), int(leaf))
        same_tree(ttree, case, "t_")
        assert counts == (len(case["m2l"]), len(case["p2p"]))
        assert np.array_equal(rows_as_pairs(mrow, mcol), case["m2l"])
        assert np.array_equal(rows_as_pairs(prow, pcol), case["p2p"])
