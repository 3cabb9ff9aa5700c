"""GPU log-kernel FMM against the oracle (lists, locals, fields) and direct sums."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relative(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def cloud(n=3000, seed=0):
    """test_fmm.py:17-27."""
    rng = np.random.default_rng(seed)
    pos = np.concatenate(
        (rng.normal([[-0.3], [0.1]], 0.1, (2, n // 2)),
         rng.normal([[0.4], [-0.2]], 0.25, (2, n - n // 2))), axis=1).clip(-0.99, 0.99)
    return pos, rng.normal(0.3, 1, n)


def config2(n, seed=7):
    rng = np.random.default_rng(seed)
    half = n // 2
    centres = rng.uniform(-0.8, 0.8, (2, 16))
    per = (n - half) // 16
    clusters = (centres[:, :, None] + rng.normal(0, 0.02, (2, 16, per))).reshape(2, -1)
    pos = np.concatenate((rng.uniform(-1, 1, (2, half)), clusters), axis=1).clip(-1, 1)
    return pos, rng.choice([-1.0, 1.0], pos.shape[1])


def directed(rows, cols):
    """(dst, src) of every row entry, sorted (a row's order is a walk order)."""
    row = np.asarray(rows.cpu()).astype(np.int64)
    col = np.asarray(cols.cpu()).astype(np.int64)[: row[-1]]
    dst = np.repeat(np.arange(len(row) - 1), np.diff(row))
    pairs = np.stack((dst, col), axis=1)
    return pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))]


def as_directed(pairs, self_once):
    pairs = np.asarray(pairs)
    back = pairs[:, ::-1]
    if self_once:
        back = back[pairs[:, 0] != pairs[:, 1]]
    both = np.concatenate((pairs, back))
    return both[np.lexsort((both[:, 1], both[:, 0]))]


@pytest.mark.parametrize("leaf", [1, 8, 64])
def test_lists_equal_reference_traversal(td3, oracle, leaf):
    pos, q = cloud(3000)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", leaf_capacity=leaf)
    ref = oracle.OracleTree(pos, q, (0, 0, 1), kernel="log", leaf_capacity=leaf)
    (mrow, mcol, _), (prow, pcol, _), counts = tree._lists(0.5)
    m2l, p2p = ref.traverse(0.5)
    assert counts == (len(m2l), len(p2p))
    np.testing.assert_array_equal(directed(mrow, mcol), as_directed(m2l, False))
    np.testing.assert_array_equal(directed(prow, pcol), as_directed(p2p, True))


def test_lists_of_big_leaves_beside_a_dense_cluster(td3, oracle):
    """Big leaves beside a dense cluster: a leaf's walk reaches thousands of the cluster's
    cells, more than the breadth-first rows kernel queues, so the rows are built again
    depth first (more rows_level launches than levels); the pair sets stay the
    reference's, and so do the fields."""
    from tree_d_3d_amd import _abi

    rng = np.random.default_rng(3)
    sparse = rng.uniform(-1, 1, (2, 60))
    cluster = np.array([[0.26], [0.26]]) + rng.normal(0, 0.004, (2, 6000))
    pos = np.concatenate((sparse, cluster), axis=1).clip(-1, 1)
    q = rng.normal(size=pos.shape[1])
    print("lists: big leaves beside a cluster", flush=True)
    _abi.timing_enable(True)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", leaf_capacity=8)
    (mrow, mcol, _), (prow, pcol, _), counts = tree._lists(0.5)
    _, launches = _abi.timing_get("rows_level")
    _abi.timing_enable(False)
    assert launches > len(tree.levels) - 1
    ref = oracle.OracleTree(pos, q, (0, 0, 1), kernel="log", leaf_capacity=8)
    m2l, p2p = ref.traverse(0.5)
    assert counts == (len(m2l), len(p2p))
    np.testing.assert_array_equal(directed(mrow, mcol), as_directed(m2l, False))
    np.testing.assert_array_equal(directed(prow, pcol), as_directed(p2p, True))
    phi, field = tree.fmm(theta=0.5, precision="fp64")
    rphi, rfield = ref.fmm(theta=0.5)
    assert relative(phi, rphi) < 1e-12 and relative(field, rfield) < 1e-12


@pytest.mark.parametrize("order", [2, 6, 12, 18])
def test_fmm_fp64_matches_oracle(td3, oracle, order):
    pos, q = cloud(3000)
    print(f"fmm order={order}", flush=True)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", order=order)
    ref = oracle.OracleTree(pos, q, (0, 0, 1), kernel="log", order=order)
    phi, field = tree.fmm(theta=0.5)
    rphi, rfield = ref.fmm(theta=0.5)
    assert tree.fmm_pairs == ref.fmm_pairs
    assert relative(phi, rphi) < 1e-12
    assert relative(field, rfield) < 1e-12
    lre, lim = tree.local_expansions(0.5)
    # compare the radius-scaled locals b_l r^l (r the cell's half-diagonal): all O(1)
    r = np.sqrt(2.0) * ref.node_half[:, None] ** np.arange(order + 1)[None, :]
    r = (np.sqrt(2.0) * ref.node_half[:, None]) ** np.arange(order + 1)[None, :]
    for got, want in ((lre, ref.local_re), (lim, ref.local_im)):
        scaled_want = want * r
        assert np.max(np.abs(got * r - scaled_want)) < 1e-11 * np.max(np.abs(scaled_want))


@pytest.mark.parametrize("order,tolerance", [(6, 1e-3), (10, 1e-5), (18, 1e-8)])
def test_fmm_converges_with_expansion_order(td3, oracle, order, tolerance):
    # test_fmm.py:38-48
    pos, q = cloud()
    ref_phi, ref_field = oracle.direct_field(pos, pos, q, exclude=np.arange(3000), kernel="log")
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", order=order)
    phi, field = tree.fmm(theta=0.5)
    assert relative(phi, ref_phi) < tolerance
    assert relative(field, ref_field) < 10 * tolerance


def test_fmm_mixed_precision_error(td3, oracle):
    pos, q = config2(200000)
    print("fmm mixed n=200000", flush=True)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", order=12)
    p64, f64 = tree.fmm(theta=0.5, softening=1e-4, precision="fp64")
    pmx, fmx = tree.fmm(theta=0.5, softening=1e-4, precision="mixed")
    assert relative(pmx, p64) < 1e-6
    assert relative(fmx, f64) < 1e-6
    idx = np.random.default_rng(0).choice(pos.shape[1], 512, replace=False)
    dphi, dfield = oracle.direct_field(pos[:, idx], pos, q, exclude=idx, softening=1e-4,
                                       kernel="log")
    assert relative(f64[:, idx], dfield) < 1e-2  # far cells are unsoftened (test_fmm.py:66-83)


def test_fmm_edges(td3):
    pos, q = cloud(400, seed=2)
    q[::7] = 0.0
    pos[:, 2] = pos[:, 1]
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log")
    tree.fmm(softening=0.01)
    with pytest.raises(ValueError, match="Coincident"):
        tree.fmm()
    with pytest.raises(ValueError, match="theta"):
        tree.fmm(theta=1.0)
    nophi, field = tree.fmm(softening=0.01, potential=False)
    assert not np.any(nophi)
    one = td3.FlatQuadtree([[0.1], [0.2]], [1.0], td3.Quad(0, 0, 1), kernel="log")
    phi, field = one.fmm()
    assert phi[0] == 0.0 and not np.any(field)


def test_fmm_net_force_is_at_the_truncation_level(td3):
    # test_fmm.py:85-91
    pos, q = cloud(6000, seed=3)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", order=18)
    _, field = tree.fmm(theta=0.5)
    force = q * field
    assert np.abs(force.sum(axis=1)).max() < 1e-11 * np.abs(force).sum()


def test_fmm_set_charges_keeps_lists(td3):
    rng = np.random.default_rng(8)
    pos = rng.uniform(-1, 1, (2, 3000))
    first, second = rng.normal(size=3000), rng.normal(size=3000)
    tree = td3.FlatQuadtree(pos, first, td3.Quad(0, 0, 1), kernel="log", leaf_capacity=16)
    tree.fmm()
    tree.set_charges(second)
    fresh = td3.FlatQuadtree(pos, second, td3.Quad(0, 0, 1), kernel="log", leaf_capacity=16)
    np.testing.assert_allclose(tree.fmm()[1], fresh.fmm()[1], rtol=0, atol=1e-10)


def test_fmm_is_deterministic(td3):
    pos, q = config2(50000, seed=3)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log")
    a = tree.fmm(softening=1e-4)
    b = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log").fmm(softening=1e-4)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("potential", [True, False])
def test_fmm_on_two_streams_equals_one_launch(td3, precision, potential):
    """fmm() runs the near field (td3_fmm_near) on a second stream beside the far field
    and adds it in the L2P (td3_fmm_far): bit for bit the one-launch td3_fmm_eval on the
    same local expansions and rows, on a fresh tree (rows built beside the moments) and
    on one whose rows are kept."""
    import torch
    from tree_d_3d_amd import _abi

    pos, q = config2(60000, seed=11)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", precision=precision)
    prec = _abi.FP64 if precision == "fp64" else _abi.MIXED
    n = pos.shape[1]
    # fresh rows with the default split of the near field, then kept rows with none of it
    # early and all of it early
    for early in (None, 0.0, 1.0):
        if early is not None:
            tree.NEAR_EARLY = early
        phi, field = tree.fmm(softening=1e-4, potential=potential)
        lre, lim = tree._local_expansions(0.5, prec)
        _, (prow, pcol, _), _ = tree._lists(0.5)
        want_phi = torch.empty(n, dtype=torch.float64, device="cuda")
        want_field = torch.empty((2, n), dtype=torch.float64, device="cuda")
        status = torch.zeros(1, dtype=torch.int64, device="cuda")
        _abi.call("td3_fmm_eval", _abi.LOG, prec, tree.order, _abi.ptr(tree._sorted_pos),
                  _abi.ptr(tree._scharge), _abi.ptr(tree._d["perm"]), n,
                  _abi.ptr(tree._d["node_center"]), _abi.ptr(tree._d["node_half"]),
                  _abi.ptr(tree._node_info), tree.node_count, _abi.ptr(lre), _abi.ptr(lim),
                  _abi.ptr(prow), _abi.ptr(pcol), _abi.ptr(tree._leaf_index()),
                  _abi.ptr(tree._sources4()), 1e-4, 1 if potential else 0,
                  _abi.ptr(want_phi), _abi.ptr(want_field), _abi.ptr(status), _abi.stream())
        assert int(status.item()) == 0
        np.testing.assert_array_equal(np.asarray(phi), want_phi.cpu().numpy())
        np.testing.assert_array_equal(np.asarray(field), want_field.cpu().numpy())
        if not potential:
            assert not np.asarray(phi).any()


def test_fp64_fmm_on_rows_of_a_mixed_fmm(td3):
    """The mixed FMM builds its rows without each M2L entry's destination (its M2L walks
    rows by their starts); an FP64 FMM on those kept rows makes the destinations first and
    gives, bit for bit, what a tree whose rows were built for FP64 gives."""
    rng = np.random.default_rng(5)
    pos, q = rng.uniform(-1, 1, (2, 80000)), rng.choice([-1.0, 1.0], 80000)
    mixed_first = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log")
    mixed_first.fmm(softening=1e-4, precision="mixed")
    assert mixed_first._lists(0.5)[0][2] is None
    got = mixed_first.fmm(softening=1e-4, precision="fp64")
    want = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log").fmm(softening=1e-4,
                                                                        precision="fp64")
    np.testing.assert_array_equal(np.asarray(got[0]), np.asarray(want[0]))
    np.testing.assert_array_equal(np.asarray(got[1]), np.asarray(want[1]))


def test_abi_stream_is_torchs_current_stream(td3):
    """_abi.stream() (torch's raw current stream) names the stream torch would use, on the
    default stream and inside a stream context."""
    import torch
    from tree_d_3d_amd import _abi

    _abi.device()
    assert _abi.stream() == torch.cuda.current_stream().cuda_stream
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        assert _abi.stream() == side.cuda_stream == torch.cuda.current_stream().cuda_stream
    assert _abi.stream() == torch.cuda.current_stream().cuda_stream


def test_moments_made_on_first_use(td3, oracle):
    """The cells' moments wait for their first reader (fmm, the arrays): the same values
    whichever comes first, and again after set_charges."""
    pos, q = cloud(2000, seed=4)
    ref = oracle.OracleTree(pos, q, (0, 0, 1), kernel="log", order=10)
    scale = np.abs(ref.moment_re).max()
    for first in ("array", "fmm"):
        tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", order=10)
        if first == "fmm":
            tree.fmm()
        np.testing.assert_allclose(np.asarray(tree.moment_re), ref.moment_re, rtol=0,
                                   atol=1e-12 * scale)
        np.testing.assert_allclose(np.asarray(tree.node_q), ref.node_q, rtol=0,
                                   atol=1e-12 * scale)
    tree.set_charges(-q)
    np.testing.assert_allclose(np.asarray(tree.moment_im), -ref.moment_im, rtol=0,
                               atol=1e-12 * scale)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fmm_at_matches_oracle(td3, oracle, precision):
    rng = np.random.default_rng(13)
    pos, q = cloud(4000, seed=13)
    targets = np.concatenate((rng.uniform(-0.9, 0.9, (2, 1500)), pos[:, :500]), axis=1)
    exclude = np.concatenate((np.full(1500, -1), np.arange(500)))
    print(f"fmm_at precision={precision}", flush=True)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", order=16, leaf_capacity=16)
    ref = oracle.OracleTree(pos, q, (0, 0, 1), kernel="log", order=16, leaf_capacity=16)
    phi, field = tree.fmm_at(targets, exclude=exclude, theta=0.5, precision=precision)
    rphi, rfield = ref.fmm_at(targets, exclude=exclude, theta=0.5)
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert relative(phi, rphi) < tol and relative(field, rfield) < tol
    assert tree.fmm_at_pairs == (len(ref.dual_lists[0]), len(ref.dual_lists[1]))
    nophi, same = tree.fmm_at(targets, exclude=exclude, theta=0.5, potential=False,
                              precision=precision)
    assert not np.any(nophi)
    np.testing.assert_allclose(same, field, rtol=1e-12, atol=1e-12)


def test_fmm_at_plan_kept_across_new_charges(td3):
    # test_fmm.py:252-270
    rng = np.random.default_rng(14)
    pos = rng.uniform(-0.9, 0.9, (2, 800))
    targets = rng.uniform(-0.9, 0.9, (2, 300))
    q, other = rng.normal(size=800), rng.normal(size=800)
    other[:400] = 0.0
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", leaf_capacity=8)
    tree.fmm_at(targets)
    tree.set_charges(other)
    fresh = td3.FlatQuadtree(pos, other, td3.Quad(0, 0, 1), kernel="log", leaf_capacity=8)
    np.testing.assert_allclose(tree.fmm_at(targets)[1], fresh.fmm_at(targets)[1], rtol=1e-12,
                               atol=1e-12)


def test_fmm_at_validation(td3):
    # test_fmm.py:272-282
    pos = np.array([[0.1, 0.5], [0.2, -0.3]])
    tree = td3.FlatQuadtree(pos, np.ones(2), td3.Quad(0, 0, 1), kernel="log")
    with pytest.raises(ValueError, match="Coincident"):
        tree.fmm_at(pos[:, :1])
    phi, field = tree.fmm_at(pos[:, :1], exclude=[0])
    _, ref = td3.direct_field(pos[:, :1], pos, np.ones(2), exclude=[0], kernel="log")
    np.testing.assert_allclose(field, ref, rtol=1e-12)
    with pytest.raises(ValueError, match="theta"):
        tree.fmm_at(pos, theta=1.0)
    assert tree.fmm_at(np.zeros((2, 0)))[1].shape == (2, 0)


@pytest.mark.parametrize("softening", [0.0, 0.02])
def test_coulomb_fmm_matches_oracle(td3, oracle, softening):
    rng = np.random.default_rng(11)
    pos = np.concatenate((rng.uniform(-0.9, 0.9, (2, 1500)), rng.normal(0.3, 0.05, (2, 1500))),
                         axis=1).clip(-0.99, 0.99)
    q = rng.normal(size=pos.shape[1])
    own = np.arange(pos.shape[1])
    _, exact = oracle.direct_field(pos, pos, q, exclude=own, softening=softening)
    errors = []
    for order in (2, 4, 6, 8, 10, 13):
        print(f"coulomb fmm order={order} eps={softening}", flush=True)
        tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), order=order, leaf_capacity=16)
        phi, field = tree.fmm(theta=0.5, softening=softening)
        ref = oracle.OracleTree(pos, q, (0, 0, 1), order=order, leaf_capacity=16)
        rphi, rfield = ref.fmm(theta=0.5, softening=softening)
        assert tree.fmm_pairs == ref.fmm_pairs
        assert relative(phi, rphi) < 1e-12 and relative(field, rfield) < 1e-11
        errors.append(relative(field, exact))
        net = np.abs((q * field).sum(axis=1)).max()
        assert net < 1e-12 * np.abs(q * field).sum()
    assert np.all(np.diff(errors[:4]) < 0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_coulomb_fmm_at_matches_oracle(td3, oracle, precision):
    rng = np.random.default_rng(13)
    pos, q = cloud(4000, seed=13)
    targets = np.concatenate((rng.uniform(-0.9, 0.9, (2, 1500)), pos[:, :500]), axis=1)
    exclude = np.concatenate((np.full(1500, -1), np.arange(500)))
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), order=8, leaf_capacity=16)
    ref = oracle.OracleTree(pos, q, (0, 0, 1), order=8, leaf_capacity=16)
    phi, field = tree.fmm_at(targets, exclude=exclude, theta=0.5, softening=0.01,
                             precision=precision)
    rphi, rfield = ref.fmm_at(targets, exclude=exclude, theta=0.5, softening=0.01)
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert relative(phi, rphi) < tol and relative(field, rfield) < tol
    # test_fmm.py:244-246: fmm_at against the walk
    _, walk = tree.evaluate(targets, exclude=exclude, theta=0.5, softening=0.01)
    _, direct = oracle.direct_field(targets, pos, q, exclude=exclude, softening=0.01)
    assert np.max(np.abs(field - walk)) < 20 * 2e-4 * np.max(np.abs(direct))


@pytest.mark.parametrize("kernel", ["log", "coulomb"])
def test_mixed_near_field_of_close_pairs_across_cells(td3, oracle, kernel):
    # pairs 1e-7 apart straddling the dyadic lines x = 0.5 and y = -0.25 (a cell boundary
    # at every depth): the mixed near field must keep their differences exact
    pos, q = cloud(3000, seed=21)
    rng = np.random.default_rng(22)
    ys = rng.uniform(-0.9, 0.9, 40)
    xs = rng.uniform(-0.9, 0.9, 40)
    close = np.concatenate((
        np.stack((np.full(40, 0.5 - 5e-8), ys)), np.stack((np.full(40, 0.5 + 5e-8), ys)),
        np.stack((xs, np.full(40, -0.25 - 5e-8))), np.stack((xs, np.full(40, -0.25 + 5e-8)))),
        axis=1)
    pos = np.concatenate((pos, close), axis=1)
    q = np.concatenate((q, rng.choice([-1.0, 1.0], close.shape[1])))
    print(f"close pairs kernel={kernel}", flush=True)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel=kernel, order=12, leaf_capacity=8)
    ref = oracle.OracleTree(pos, q, (0, 0, 1), kernel=kernel, order=12, leaf_capacity=8)
    eps = 1e-9 if kernel == "coulomb" else 0.0
    want = ref.fmm(theta=0.5, softening=eps)
    got = tree.fmm(theta=0.5, softening=eps, precision="mixed")
    assert relative(got[1], want[1]) < 1e-6 and relative(got[0], want[0]) < 1e-6
