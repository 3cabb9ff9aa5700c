"""Repeated sums over one geometry (electrodes.py:1205-1235): the tree and the target plan
are built once and kept across new charges, and every product equals the oracle's FMM (or
walk) of the same charges."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def relative(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))


def panels(n, seed):
    """Points along two closed curves and their mirror images below y = 0 (the shape of
    the reference's conductor panels with image charges)."""
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0, 2 * np.pi, n // 4))
    a = np.stack((-0.4 + 0.25 * np.cos(t), 0.45 + 0.15 * np.sin(t)))
    b = np.stack((0.35 + 0.2 * np.cos(3 * t) * np.cos(t), 0.5 + 0.2 * np.sin(t)))
    upper = np.concatenate((a, b), axis=1)
    return np.ascontiguousarray(np.concatenate((upper, upper * [[1.0], [-1.0]]), axis=1))


def test_products_reuse_the_plan_and_match_the_oracle(td3, oracle):
    pts = panels(8000, 1)
    n = pts.shape[1]
    own = np.arange(0, n // 2, 2)  # every other upper point is a target (its own excluded)
    targets = np.ascontiguousarray(pts[:, own])
    sums = td3.RepeatedSums(pts, targets, td3.Quad(0, 0, 1), exclude=own)
    assert not sums.walks
    rng = np.random.default_rng(2)
    ref = None
    print("repeated sums: 5 products, order 32", flush=True)
    for k in range(5):
        q = rng.normal(size=n)
        phi, field = sums(q)
        if ref is None:
            ref = oracle.OracleTree(pts, q, (0, 0, 1), kernel="log", order=32, leaf_capacity=16)
        else:
            ref.set_charges(q)
        rphi, rfield = ref.fmm_at(targets, exclude=own, theta=0.5)
        assert relative(phi, rphi) < 1e-12 and relative(field, rfield) < 1e-12, k
        # a fresh tree and plan give the same numbers, bit for bit
        fresh = td3.FlatQuadtree(pts, q, td3.Quad(0, 0, 1), kernel="log", order=32,
                                 leaf_capacity=16)
        fphi, ffield = fresh.fmm_at(targets, exclude=own, theta=0.5)
        assert np.array_equal(phi, fphi) and np.array_equal(field, ffield), k
    assert sums.tree_builds == 1 and sums.plan_builds == 1 and sums.products == 5
    # the SciPy operator applies the same product
    op = sums.as_linear_operator()
    q = rng.normal(size=n)
    assert np.array_equal(op.matvec(q), sums(q)[0])
    assert sums.plan_builds == 1


def test_few_targets_walk(td3, oracle):
    """electrodes.py:267-271: fewer targets than the sources over 8 take the walk at
    theta 0.7."""
    pts = panels(4000, 3)
    n = pts.shape[1]
    own = np.arange(0, n, 20)
    sums = td3.RepeatedSums(pts, pts[:, own], td3.Quad(0, 0, 1), exclude=own)
    assert sums.walks
    rng = np.random.default_rng(4)
    ref = None
    print("repeated sums: walk branch", flush=True)
    for k in range(3):
        q = rng.normal(size=n)
        phi, field = sums(q)
        if ref is None:
            ref = oracle.OracleTree(pts, q, (0, 0, 1), kernel="log", order=32, leaf_capacity=16)
        else:
            ref.set_charges(q)
        rphi, rfield = ref.evaluate(pts[:, own], exclude=own, theta=0.7)
        assert relative(phi, rphi) < 1e-12 and relative(field, rfield) < 1e-12, k
    assert sums.tree_builds == 1 and sums.plan_builds == 0


def test_torch_charges_stay_on_the_device(td3):
    pts = torch.tensor(panels(2000, 5), device="cuda")
    sums = td3.RepeatedSums(pts, pts[:, ::2].contiguous(), td3.Quad(0, 0, 1),
                            exclude=torch.arange(0, 2000, 2, device="cuda"))
    phi, field = sums(torch.ones(2000, dtype=torch.float64, device="cuda"))
    assert phi.is_cuda and field.is_cuda and phi.shape == (1000,)
    with pytest.raises(ValueError):
        sums(torch.ones(10, dtype=torch.float64, device="cuda"))
