"""The smooth-membership FMM (smooth.py) on the GPU: cells, memberships, lists, moments,
locals and fields against the oracle's restatement; the reference's own smooth tests
(test_smooth.py) run on the GPU classes; and size-independent properties at 200k particles
(partition of unity, exact antisymmetry, accuracy against the direct sum)."""

from collections import namedtuple

import numpy as np
import pytest
import torch

from oracle import smooth_oracle as so

pytestmark = pytest.mark.gpu

Box = namedtuple("Box", "xmid ymid half")


def cloud(n, seed=0, box=0.9):
    """test_smooth.py:18-20."""
    rng = np.random.default_rng(seed)
    return rng.uniform(-box, box, (2, n)), rng.choice([-1.0, 1.0], n)


def relative(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))


def net_force(q, field):
    return np.abs((q * field).sum(axis=1)).max() / np.abs(q * field).sum()


def edge_cloud(n, seed):
    pos, q = cloud(n, seed)
    pos[0, :20] = np.round(pos[0, :20] * 8) / 8  # on vertical split lines
    pos[1, 20:40] = np.round(pos[1, 20:40] * 8) / 8  # on horizontal ones
    return pos, q


def test_box_tree_matches_oracle(td3):
    pos, _ = edge_cloud(3000, 1)
    print("smooth box tree n=3000", flush=True)
    for cap in (4, 16):
        tree = td3.smooth.BoxTree(pos, td3.Quad(0.0, 0.0, 1.0), leaf_capacity=cap)
        ref = so.BoxTree(pos, Box(0.0, 0.0, 1.0), leaf_capacity=cap)
        for name in ("center", "half", "child", "parent", "depth", "leaves"):
            assert np.array_equal(getattr(tree, name), getattr(ref, name)), name
        assert tree.min_leaf_half == ref.min_leaf_half
    # torch inputs give device arrays
    t = td3.smooth.BoxTree(torch.tensor(pos, device="cuda"), td3.Quad(0.0, 0.0, 1.0))
    assert t.center.is_cuda and t.center.dtype == torch.complex128


@pytest.mark.parametrize("frac", [0.0, 0.25, 0.5, 1.0])
def test_memberships_match_oracle(td3, frac):
    pos, _ = edge_cloud(4000, 2)
    tree = td3.smooth.BoxTree(pos, td3.Quad(0.0, 0.0, 1.0), leaf_capacity=8)
    ref = so.BoxTree(pos, Box(0.0, 0.0, 1.0), leaf_capacity=8)
    z = pos[0] + 1j * pos[1]
    # memberships of other points than the build's (the cells are static)
    z2 = z + 0.003 * (np.cos(np.arange(4000)) + 1j * np.sin(np.arange(4000)))
    print(f"smooth memberships band={frac}", flush=True)
    for zz in (z, z2):
        got = tree.memberships(zz, frac * tree.min_leaf_half)
        want = ref.memberships(zz, frac * ref.min_leaf_half)
        order = np.lexsort((want[1], want[0]))
        for g, w in zip(got, want):
            assert np.array_equal(g, w[order])  # the same memberships, bit for bit
    with pytest.raises(ValueError, match="band"):
        tree.memberships(z, 1.01 * tree.min_leaf_half)


@pytest.mark.parametrize("order,symmetric,softening,frac",
                         [(8, True, 0.0, 0.0), (8, False, 0.02, 0.4), (12, True, 0.01, 0.5),
                          (20, True, 0.0, 0.3), (20, False, 0.05, 1.0)])
def test_fmm_matches_oracle(td3, order, symmetric, softening, frac):
    pos, q = edge_cloud(1200, 3)
    q = q * np.random.default_rng(4).uniform(0.5, 1.5, 1200)
    tree = td3.smooth.BoxTree(pos, td3.Quad(0.0, 0.0, 1.0), leaf_capacity=12)
    ref_tree = so.BoxTree(pos, Box(0.0, 0.0, 1.0), leaf_capacity=12)
    kw = dict(order=order, theta=0.6, band=frac * tree.min_leaf_half, symmetric=symmetric,
              softening=softening)
    print(f"smooth fmm {kw}", flush=True)
    f = td3.smooth.SmoothFMM(tree, pos, q, **kw)
    r = so.SmoothFMM(ref_tree, pos, q, **kw)
    assert sorted(map(tuple, np.sort(f.m2l_pairs, axis=1).tolist())) == \
        sorted(map(tuple, np.sort(r.m2l_pairs, axis=1).tolist()))
    assert sorted(map(tuple, np.sort(f.p2p_pairs, axis=1).tolist())) == \
        sorted(map(tuple, np.sort(r.p2p_pairs, axis=1).tolist()))
    occupied = np.abs(r.moments).sum(axis=1) > 0
    assert relative(f.moments[occupied], r.moments[occupied]) < 1e-12
    assert relative(f.local, r.local) < 1e-11
    key = np.lexsort((f.leaf, f.part))
    rkey = np.lexsort((r.leaf, r.part))
    assert np.array_equal(f.part[key], r.part[rkey]) and np.array_equal(f.w[key], r.w[rkey])
    assert relative(f.phi_m[key], r.phi_m[rkey]) < 1e-12
    assert relative(f.g_m[key], r.g_m[rkey]) < 1e-12
    phi, field = f.fields()
    rphi, rfield = r.fields()
    assert relative(phi, rphi) < 1e-12 and relative(field, rfield) < 1e-12
    scale = np.sum(np.abs(q * rphi))
    assert abs(f.energy - r.energy) < 1e-13 * scale


def test_direct_sums_match_oracle(td3):
    pos, q = cloud(900, 5)
    probes = np.random.default_rng(6).uniform(-1, 1, (2, 70))
    tree = td3.smooth.BoxTree(pos, td3.Quad(0.0, 0.0, 1.0))
    for eps in (0.0, 0.03):
        print(f"smooth direct eps={eps}", flush=True)
        phi, field = td3.smooth.direct_fields(pos, q, eps)
        rphi, rfield = so.direct_fields(pos, q, eps)
        assert relative(phi, rphi) < 1e-13 and relative(field, rfield) < 1e-13
        u, ru = td3.smooth.direct_energy(pos, q, eps), so.direct_energy(pos, q, eps)
        assert abs(u - ru) < 1e-13 * np.sum(np.abs(q * rphi))
        got = td3.smooth.SmoothFMM(tree, pos, q, order=4, softening=eps).potential_at(probes)
        want = so.SmoothFMM(so.BoxTree(pos, Box(0.0, 0.0, 1.0)), pos, q, order=4,
                            softening=eps).potential_at(probes)
        assert relative(got, want) < 1e-13
    with pytest.raises(ValueError, match="coincide"):
        td3.smooth.direct_fields([[0.1, 0.1], [0.2, 0.2]], [1.0, 1.0])
    phi, _ = td3.smooth.direct_fields([[0.1, 0.1], [0.2, 0.2]], [1.0, 1.0], 0.1)
    assert np.all(np.isfinite(phi))


# ------------------------------------------------------------------ the reference's tests
def test_partition_of_unity_and_memberships(td3):
    """test_smooth.py:40-49."""
    pos, q = cloud(400)
    tree = td3.smooth.BoxTree(pos, td3.Quad(0, 0, 1), leaf_capacity=8)
    for frac in (0.1, 0.5):
        f = td3.smooth.SmoothFMM(tree, pos, q, band=frac * tree.min_leaf_half, order=4)
        total = np.bincount(f.part, f.w, minlength=400)
        np.testing.assert_allclose(total, 1.0, atol=1e-14)
        assert 1.0 < len(f.part) / 400 < 3.0
    with pytest.raises(ValueError, match="band"):
        td3.smooth.SmoothFMM(tree, pos, q, band=1.01 * tree.min_leaf_half, order=4)


def test_sharp_membership_is_the_ordinary_tree(td3):
    """test_smooth.py:52-60."""
    pos, q = cloud(600)
    tree = td3.smooth.BoxTree(pos, td3.Quad(0, 0, 1), leaf_capacity=16)
    f = td3.smooth.SmoothFMM(tree, pos, q, order=12, theta=0.5, band=0.0)
    assert len(f.part) == 600 and np.all(f.w == 1.0)
    _, field = f.fields()
    _, ref_field = td3.smooth.direct_fields(pos, q)
    assert np.abs(field - ref_field).max() < 1e-6 * np.abs(ref_field).max()
    u = td3.smooth.direct_energy(pos, q)
    assert abs(f.energy - u) < 1e-7 * abs(u)


@pytest.mark.parametrize("symmetric", [True, False])
def test_forces_are_the_exact_gradient_of_the_energy(td3, symmetric):
    """test_smooth.py:63-86."""
    pos, q = cloud(300, seed=2)
    tree = td3.smooth.BoxTree(pos, td3.Quad(0, 0, 1), leaf_capacity=8)
    kw = dict(order=8, theta=0.6, band=0.4 * tree.min_leaf_half, softening=0.02,
              symmetric=symmetric)
    _, field = td3.smooth.SmoothFMM(tree, pos, q, **kw).fields()
    force = q * field
    f = td3.smooth.SmoothFMM(tree, pos, q, **kw)
    counts = np.bincount(f.part, minlength=300)
    rng = np.random.default_rng(5)
    blended = rng.choice(np.flatnonzero(counts > 1), 4, replace=False)
    single = rng.choice(np.flatnonzero(counts == 1), 2, replace=False)
    scale = np.abs(force).max()
    h = 1e-6
    for i in [*blended, *single]:
        for axis in range(2):
            up, down = pos.copy(), pos.copy()
            up[axis, i] += h
            down[axis, i] -= h
            fd = -(td3.smooth.SmoothFMM(tree, up, q, **kw).energy
                   - td3.smooth.SmoothFMM(tree, down, q, **kw).energy) / (2 * h)
            assert abs(fd - force[axis, i]) < 1e-7 * scale


def test_symmetric_truncation_gives_exact_antisymmetry(td3):
    """test_smooth.py:89-104."""
    pos, q = cloud(1500, seed=1)
    tree = td3.smooth.BoxTree(pos, td3.Quad(0, 0, 1), leaf_capacity=16)
    ref = td3.smooth.direct_fields(pos, q)[1]
    for order in (8, 16):
        _, e_rect = td3.smooth.SmoothFMM(tree, pos, q, order=order, theta=0.5,
                                         symmetric=False).fields()
        _, e_tri = td3.smooth.SmoothFMM(tree, pos, q, order=order, theta=0.5,
                                        symmetric=True).fields()
        assert net_force(q, e_tri) < 1e-15
        assert net_force(q, e_rect) > 1e-13
        err_rect = np.abs(e_rect - ref).max() / np.abs(ref).max()
        err_tri = np.abs(e_tri - ref).max() / np.abs(ref).max()
        assert err_rect < err_tri < 100 * err_rect


def test_smooth_membership_keeps_the_accuracy(td3):
    """test_smooth.py:107-119."""
    pos, q = cloud(800, seed=4)
    tree = td3.smooth.BoxTree(pos, td3.Quad(0, 0, 1), leaf_capacity=8)
    ref = td3.smooth.direct_fields(pos, q)[1]
    errors = []
    for frac in (0.0, 0.25, 0.5):
        _, e = td3.smooth.SmoothFMM(tree, pos, q, order=12, theta=0.5,
                                    band=frac * tree.min_leaf_half).fields()
        errors.append(np.abs(e - ref).max() / np.abs(ref).max())
    assert errors[0] < 1e-6
    assert errors[2] < 50 * errors[0]


def _verlet(pos, vel, q, m, dt, steps, k_trap, fields):
    """test_smooth.py:122-144."""
    x, v = pos.copy(), vel.copy()
    u, e = fields(x)
    e0 = u + 0.5 * k_trap * np.sum(x ** 2) + 0.5 * np.sum(m * np.sum(v ** 2, axis=0))
    acc = (q * e - k_trap * x) / m
    drift, net = [], 0.0
    for _ in range(steps):
        v += 0.5 * dt * acc
        x += dt * v
        u, e = fields(x)
        acc = (q * e - k_trap * x) / m
        v += 0.5 * dt * acc
        drift.append((u + 0.5 * k_trap * np.sum(x ** 2) + 0.5 * np.sum(m * np.sum(v ** 2, axis=0)))
                     / e0 - 1.0)
        net = max(net, np.abs((q * e).sum(axis=1)).max() / np.abs(q * e).sum())
    return np.array(drift), net


def test_comoving_boxes_conserve_energy_and_momentum(td3):
    """test_smooth.py:147-186."""
    n, steps, dt, k_trap, eps = 150, 300, 0.002, 100.0, 0.1
    rng = np.random.default_rng(3)
    pos = rng.uniform(-0.8, 0.8, (2, n))
    q = rng.choice([-1.0, 1.0], n)
    vel = 0.2 * rng.normal(size=(2, n))
    m = np.ones(n)
    tree = td3.smooth.BoxTree(pos, td3.Quad(0, 0, 2.0), leaf_capacity=8)
    center = (m * pos).sum(axis=1) / m.sum()
    options = dict(order=2, theta=0.8, softening=eps, symmetric=True)

    def direct(x):
        return td3.smooth.direct_energy(x, q, eps), td3.smooth.direct_fields(x, q, eps)[1]

    def tree_fields(frac):
        def fields(x):
            u, _, e = td3.smooth.comoving_fields(tree, x, q, m, center,
                                                 band=frac * tree.min_leaf_half, **options)
            return u, e
        return fields

    print("smooth comoving verlet", flush=True)
    reference, _ = _verlet(pos, vel, q, m, dt, steps, k_trap, direct)
    sharp, net_sharp = _verlet(pos, vel, q, m, dt, steps, k_trap, tree_fields(0.0))
    smooth, net_smooth = _verlet(pos, vel, q, m, dt, steps, k_trap, tree_fields(0.5))
    assert net_sharp < 1e-14 and net_smooth < 1e-14
    assert np.abs(smooth).max() < 1.5 * np.abs(reference).max()
    assert np.abs(sharp).max() > 5.0 * np.abs(smooth).max()
    # the comoving field against the oracle's on one configuration
    _, phi, field = td3.smooth.comoving_fields(tree, pos, q, m, center, band=0.5 * tree.min_leaf_half,
                                               **options)
    _, rphi, rfield = so.comoving_fields(so.BoxTree(pos, Box(0, 0, 2.0), leaf_capacity=8), pos, q, m,
                                         center, band=0.5 * tree.min_leaf_half, **options)
    assert relative(phi, rphi) < 1e-12 and relative(field, rfield) < 1e-12


# ------------------------------------------------------------------ at scale and edges
def test_large_cloud_properties(td3):
    """200k particles: with the band half the smallest leaf the weights sum to one and the
    field agrees with the direct sum; with hard membership the mutual forces cancel exactly
    under the triangle truncation (with a band, the cells stay where they are while the
    particles move, so the energy is not translation invariant and the net force is not
    zero: comoving_fields is what restores it)."""
    n = 200_000
    rng = np.random.default_rng(9)
    pos = torch.tensor(np.concatenate((rng.normal(0.0, 0.25, (2, n // 2)).clip(-0.99, 0.99),
                                       rng.uniform(-0.99, 0.99, (2, n - n // 2))), axis=1),
                       device="cuda")
    q = torch.tensor(rng.choice([-1.0, 1.0], n), device="cuda")
    print("smooth large cloud n=200000", flush=True)
    tree = td3.smooth.BoxTree(pos, td3.Quad(0.0, 0.0, 1.0), leaf_capacity=16)
    f = td3.smooth.SmoothFMM(tree, pos, q, order=16, theta=0.5, band=0.5 * tree.min_leaf_half,
                             softening=1e-3)
    total = torch.zeros(n, dtype=torch.float64, device="cuda").index_add_(0, f.part, f.w)
    assert float((total - 1).abs().max()) < 1e-13
    phi, field = f.fields()
    _, hard = td3.smooth.SmoothFMM(tree, pos, q, order=16, theta=0.5, softening=1e-3).fields()
    force = q * hard
    assert float(force.sum(dim=1).abs().max() / force.abs().sum()) < 1e-15
    ref_phi, ref_field = td3.smooth.direct_fields(pos, q, 1e-3)
    assert float((field - ref_field).abs().max() / ref_field.abs().max()) < 1e-5
    u = td3.smooth.direct_energy(pos, q, 1e-3)
    assert abs(f.energy - u) < 1e-6 * float((q * ref_phi).abs().sum())


def test_edges_and_validation(td3):
    empty = td3.smooth.BoxTree(np.zeros((2, 0)), td3.Quad(0, 0, 1))
    assert empty.child.tolist() == [-1]
    f = td3.smooth.SmoothFMM(empty, np.zeros((2, 0)), [])
    phi, field = f.fields()
    assert f.energy == 0.0 and phi.shape == (0,) and field.shape == (2, 0)
    lone = td3.smooth.SmoothFMM(empty, [[0.2], [0.3]], [2.0], band=0.5)
    assert lone.energy == 0.0 and np.all(lone.fields()[1] == 0.0)
    pos, q = cloud(50, 3)
    tree = td3.smooth.BoxTree(pos, td3.Quad(0, 0, 1), leaf_capacity=4)
    with pytest.raises(ValueError, match="order"):
        td3.smooth.SmoothFMM(tree, pos, q, order=0)
    with pytest.raises(ValueError, match="same length"):
        td3.smooth.SmoothFMM(tree, pos, q[:-1])
    with pytest.raises(ValueError, match="coincide"):
        td3.smooth.SmoothFMM(tree, np.concatenate((pos, pos[:, :1]), axis=1),
                             np.concatenate((q, q[:1])))
    both = td3.smooth.SmoothFMM(tree, np.concatenate((pos, pos[:, :1]), axis=1),
                                np.concatenate((q, q[:1])), softening=0.01)
    assert np.isfinite(both.energy)
