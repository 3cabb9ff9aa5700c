"""Pins the oracle's smooth-membership FMM (oracle/smooth_oracle.py, restating smooth.py)
by what fixes it independently of its own formulas: the reference's own tests
(test_smooth.py) run against it, the compact softening against Gauss's law, the
memberships against a scan over every leaf, and the interaction lists against the rule that
every pair of occupied leaves is served exactly once."""

from collections import namedtuple

import numpy as np
import pytest
from scipy.integrate import quad

from oracle import smooth_oracle as so

Quad = namedtuple("Quad", "xmid ymid half")


def cloud(n, seed=0, box=0.9):
    """test_smooth.py:18-20."""
    rng = np.random.default_rng(seed)
    return rng.uniform(-box, box, (2, n)), rng.choice([-1.0, 1.0], n)


def net_force(q, field):
    return np.abs((q * field).sum(axis=1)).max() / np.abs(q * field).sum()


def fd_force(tree, pos, q, kw, indices, h=1e-6):
    """test_smooth.py:27-37: -dU/dx by central differences."""
    out = {}
    for i in indices:
        for axis in range(2):
            up, down = pos.copy(), pos.copy()
            up[axis, i] += h
            down[axis, i] -= h
            out[i, axis] = -(so.SmoothFMM(tree, up, q, **kw).energy
                             - so.SmoothFMM(tree, down, q, **kw).energy) / (2 * h)
    return out


def test_partition_of_unity_and_band_limit():
    """test_smooth.py:40-49."""
    pos, q = cloud(400)
    tree = so.BoxTree(pos, Quad(0, 0, 1), leaf_capacity=8)
    for frac in (0.1, 0.5, 1.0):
        f = so.SmoothFMM(tree, pos, q, band=frac * tree.min_leaf_half, order=4)
        np.testing.assert_allclose(np.bincount(f.part, f.w, minlength=400), 1.0, atol=1e-14)
        assert 1.0 < len(f.part) / 400 < 3.0
        assert np.bincount(f.part, minlength=400).max() <= 4
    with pytest.raises(ValueError, match="band"):
        so.SmoothFMM(tree, pos, q, band=1.01 * tree.min_leaf_half, order=4)


def test_memberships_against_every_leaf():
    """The descent keeps a leaf exactly when the point lies strictly inside the leaf grown
    by the band; the weight is the product of the four edge steps, each edge on the root's
    boundary at 1; the gradient is the weight's (central differences)."""
    pos, _ = cloud(300, seed=3, box=1.0)
    pos[:, :20] = np.clip(pos[:, :20] * 1.5, -1.0, 1.0)  # some on and near the root edge
    tree = so.BoxTree(pos, Quad(0.0, 0.0, 1.0), leaf_capacity=6)
    band = 0.7 * tree.min_leaf_half
    z = pos[0] + 1j * pos[1]
    part, leaf, w, dwx, dwy = tree.memberships(z, band)
    got = {(int(p), int(l)): (a, b, c) for p, l, a, b, c in zip(part, leaf, w, dwx, dwy)}

    def weight(x, y, cell):
        c, h = tree.center[cell], tree.half[cell]
        total = 1.0
        for dist, edge, root_edge in ((x - (c.real - h), c.real - h, -1.0),
                                      ((c.real + h) - x, c.real + h, 1.0),
                                      (y - (c.imag - h), c.imag - h, -1.0),
                                      ((c.imag + h) - y, c.imag + h, 1.0)):
            on_root = abs(edge - root_edge) < 1e-12
            u = min(max((dist / band + 1.0) / 2.0, 0.0), 1.0)
            total *= 1.0 if on_root else u ** 3 * (6 * u * u - 15 * u + 10)
        return total

    want = {}
    for i in range(pos.shape[1]):
        for cell in tree.leaves:
            c, h = tree.center[cell], tree.half[cell]
            if abs(pos[0, i] - c.real) < h + band and abs(pos[1, i] - c.imag) < h + band:
                wv = weight(pos[0, i], pos[1, i], cell)
                if wv > 0:
                    want[i, int(cell)] = wv
    assert set(got) == set(want)
    hstep = 1e-7
    for (i, cell), wv in want.items():
        assert got[i, cell][0] == pytest.approx(wv, rel=1e-12, abs=1e-15)
        gx = (weight(pos[0, i] + hstep, pos[1, i], cell)
              - weight(pos[0, i] - hstep, pos[1, i], cell)) / (2 * hstep)
        gy = (weight(pos[0, i], pos[1, i] + hstep, cell)
              - weight(pos[0, i], pos[1, i] - hstep, cell)) / (2 * hstep)
        assert abs(got[i, cell][1] - gx) < 1e-5 / band and abs(got[i, cell][2] - gy) < 1e-5 / band


def test_box_tree_cells_and_hard_leaves():
    """Cells hold at most leaf_capacity build points; y == cy counts as below the centre
    (smooth.py:110, :135-137) and x == cx as left."""
    pos, _ = cloud(500, seed=5)
    pos[:, :40] = np.round(pos[:, :40] * 4) / 4  # points on cell edges
    tree = so.BoxTree(pos, Quad(0.0, 0.0, 1.0), leaf_capacity=10)
    z = pos[0] + 1j * pos[1]
    leaf = tree.hard_leaf(z)
    assert np.bincount(leaf, minlength=len(tree.half)).max() <= 10
    for i in range(pos.shape[1]):
        c, h = tree.center[leaf[i]], tree.half[leaf[i]]
        assert c.real - h <= z[i].real <= c.real + h and c.imag - h <= z[i].imag <= c.imag + h
        # a point on a split line belongs to the cell left of it and the cell below it, so
        # it never lies on its leaf's left or bottom edge (unless that is the root's)
        assert not (z[i].real == c.real - h and c.real - h > -1.0)
        assert not (z[i].imag == c.imag - h and c.imag - h > -1.0)
    assert any(z[i].imag == tree.center[leaf[i]].imag + tree.half[leaf[i]] < 1.0
               for i in range(40))
    # band 0: one membership per particle, weight 1
    part, lf, w, dwx, dwy = tree.memberships(z, 0.0)
    assert np.array_equal(part, np.arange(500)) and np.array_equal(lf, leaf)
    assert np.all(w == 1.0) and not dwx.any() and not dwy.any()


def test_lists_serve_every_pair_of_leaves_once():
    """Every ordered pair of occupied leaves is served by exactly one M2L pair of their
    ancestors-or-selves (either way round) or one P2P pair; M2L pairs are well separated
    and P2P pairs of distinct leaves are not."""
    pos, q = cloud(700, seed=6)
    tree = so.BoxTree(pos, Quad(0, 0, 1), leaf_capacity=8)
    for theta, band in ((0.5, 0.0), (0.6, 0.4 * tree.min_leaf_half)):
        f = so.SmoothFMM(tree, pos, q, order=3, theta=theta, band=band)
        occupied_leaves = np.unique(f.leaf)
        anc = {}
        for L in occupied_leaves:
            chain, c = [], int(L)
            while c >= 0:
                chain.append(c)
                c = int(tree.parent[c])
            anc[int(L)] = chain
        count = {(int(a), int(b)): 0 for a in occupied_leaves for b in occupied_leaves}
        where = {}
        for a in occupied_leaves:
            for c in anc[int(a)]:
                where.setdefault(c, []).append(int(a))
        for a, b in f.m2l_pairs:
            assert np.sqrt(2) * (tree.half[a] + tree.half[b]) < theta * abs(
                tree.center[a] - tree.center[b])
            for la in where.get(int(a), []):
                for lb in where.get(int(b), []):
                    count[la, lb] += 1
                    count[lb, la] += 1
        for a, b in f.p2p_pairs:
            assert tree.child[a] < 0 and tree.child[b] < 0
            if a != b:
                assert not np.sqrt(2) * (tree.half[a] + tree.half[b]) < theta * abs(
                    tree.center[a] - tree.center[b])
                count[int(b), int(a)] += 1
            count[int(a), int(b)] += 1
        assert set(count.values()) == {1}


def test_sharp_membership_is_the_ordinary_tree():
    """test_smooth.py:52-60."""
    pos, q = cloud(600)
    tree = so.BoxTree(pos, Quad(0, 0, 1), leaf_capacity=16)
    f = so.SmoothFMM(tree, pos, q, order=12, theta=0.5, band=0.0)
    assert len(f.part) == 600 and np.all(f.w == 1.0)
    _, field = f.fields()
    _, ref_field = so.direct_fields(pos, q)
    assert np.abs(field - ref_field).max() < 1e-6 * np.abs(ref_field).max()
    ref_u = so.direct_energy(pos, q)
    assert abs(f.energy - ref_u) < 1e-7 * abs(ref_u)


@pytest.mark.parametrize("symmetric", [True, False])
def test_forces_are_the_exact_gradient_of_the_energy(symmetric):
    """test_smooth.py:63-86."""
    pos, q = cloud(300, seed=2)
    tree = so.BoxTree(pos, Quad(0, 0, 1), leaf_capacity=8)
    kw = dict(order=8, theta=0.6, band=0.4 * tree.min_leaf_half, softening=0.02,
              symmetric=symmetric)
    f = so.SmoothFMM(tree, pos, q, **kw)
    _, field = f.fields()
    force = q * field
    counts = np.bincount(f.part, minlength=300)
    rng = np.random.default_rng(5)
    blended = rng.choice(np.flatnonzero(counts > 1), 4, replace=False)
    single = rng.choice(np.flatnonzero(counts == 1), 2, replace=False)
    scale = np.abs(force).max()
    for (i, axis), value in fd_force(tree, pos, q, kw, [*blended, *single]).items():
        assert abs(value - force[axis, i]) < 1e-7 * scale


def test_symmetric_truncation_gives_exact_antisymmetry():
    """test_smooth.py:89-104."""
    pos, q = cloud(1500, seed=1)
    tree = so.BoxTree(pos, Quad(0, 0, 1), leaf_capacity=16)
    ref = so.direct_fields(pos, q)[1]
    for order in (8, 16):
        _, e_rect = so.SmoothFMM(tree, pos, q, order=order, theta=0.5, symmetric=False).fields()
        _, e_tri = so.SmoothFMM(tree, pos, q, order=order, theta=0.5, symmetric=True).fields()
        assert net_force(q, e_tri) < 1e-15
        assert net_force(q, e_rect) > 1e-13
        err_rect = np.abs(e_rect - ref).max() / np.abs(ref).max()
        err_tri = np.abs(e_tri - ref).max() / np.abs(ref).max()
        assert err_rect < err_tri < 100 * err_rect


def test_smooth_membership_keeps_the_accuracy():
    """test_smooth.py:107-119."""
    pos, q = cloud(800, seed=4)
    tree = so.BoxTree(pos, Quad(0, 0, 1), leaf_capacity=8)
    ref = so.direct_fields(pos, q)[1]
    errors = []
    for frac in (0.0, 0.25, 0.5):
        _, e = so.SmoothFMM(tree, pos, q, order=12, theta=0.5,
                            band=frac * tree.min_leaf_half).fields()
        errors.append(np.abs(e - ref).max() / np.abs(ref).max())
    assert errors[0] < 1e-6
    assert errors[2] < 50 * errors[0]


def test_compact_softening_is_gauss_law():
    """smooth.py:204-221: the pair energy of two unit charges inside the cloud is -2 times
    ln a minus the integral of the enclosed fraction, here by quadrature and by hand; the
    kernel is the bare ln r outside, continuous at the rim, and its gradient is the
    derivative of its potential."""
    rng = np.random.default_rng(7)
    for _ in range(8):
        r, a = rng.uniform(0.05, 0.5), rng.uniform(0.6, 1.0)
        integral = quad(lambda s: (1.0 - (1.0 - s * s / (a * a)) ** 3) / s, r, a,
                        epsabs=1e-14, epsrel=1e-13)[0]
        want = np.log(a) - integral
        got = -0.5 * so.direct_energy(np.array([[0.0, r], [0.0, 0.0]]), np.ones(2), softening=a)
        assert got == pytest.approx(want, rel=1e-12, abs=1e-13)
        assert so.softening_closed_form(r, a) == pytest.approx(want, rel=1e-12, abs=1e-13)
    a = 0.3
    for r in (a * (1 - 1e-9), a * (1 + 1e-9)):
        phi, _ = so.compact_log(np.array([r + 0j]), a * a)
        assert phi[0] == pytest.approx(np.log(a), abs=1e-8)
    for dz in (0.1 + 0.05j, 0.2 - 0.15j, 0.5 + 0.1j):
        h = 1e-6
        _, g = so.compact_log(np.array([dz]), a * a)
        gx = (so.compact_log(np.array([dz + h]), a * a)[0] - so.compact_log(np.array([dz - h]), a * a)[0]) / (2 * h)
        gy = (so.compact_log(np.array([dz + 1j * h]), a * a)[0]
              - so.compact_log(np.array([dz - 1j * h]), a * a)[0]) / (2 * h)
        assert g[0].real == pytest.approx(gx[0], rel=1e-7) and g[0].imag == pytest.approx(gy[0], rel=1e-7)


def test_direct_sums_agree():
    """direct_energy is half the charge-weighted direct potential; potential_at at the
    particles themselves (self term dropped by hand) is the direct potential."""
    pos, q = cloud(120, seed=8)
    for eps in (0.0, 0.05):
        phi, _ = so.direct_fields(pos, q, eps)
        assert so.direct_energy(pos, q, eps) == pytest.approx(0.5 * np.dot(q, phi), rel=1e-12)
        tree = so.BoxTree(pos, Quad(0, 0, 1), leaf_capacity=8)
        f = so.SmoothFMM(tree, pos, q, order=4, softening=eps)
        shifted = pos + 1e-3
        want = np.array([-2.0 * np.dot(so.compact_log((shifted[0, i] - pos[0]) + 1j * (shifted[1, i] - pos[1]),
                                                     eps * eps)[0], q) for i in range(120)])
        np.testing.assert_allclose(f.potential_at(shifted), want, rtol=1e-13)


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


def test_comoving_boxes_conserve_energy_and_momentum():
    """test_smooth.py:147-186."""
    n, steps, dt, k_trap, eps = 150, 300, 0.002, 100.0, 0.1
    rng = np.random.default_rng(3)
    pos = rng.uniform(-0.8, 0.8, (2, n))
    q = rng.choice([-1.0, 1.0], n)
    vel = 0.2 * rng.normal(size=(2, n))
    m = np.ones(n)
    tree = so.BoxTree(pos, Quad(0, 0, 2.0), leaf_capacity=8)
    center = (m * pos).sum(axis=1) / m.sum()
    options = dict(order=2, theta=0.8, softening=eps, symmetric=True)

    def direct(x):
        return so.direct_energy(x, q, eps), so.direct_fields(x, q, eps)[1]

    def tree_fields(frac):
        def fields(x):
            u, _, e = so.comoving_fields(tree, x, q, m, center,
                                         band=frac * tree.min_leaf_half, **options)
            return u, e
        return fields

    reference, _ = _verlet(pos, vel, q, m, dt, steps, k_trap, direct)
    sharp, net_sharp = _verlet(pos, vel, q, m, dt, steps, k_trap, tree_fields(0.0))
    smooth, net_smooth = _verlet(pos, vel, q, m, dt, steps, k_trap, tree_fields(0.5))
    assert net_sharp < 1e-14 and net_smooth < 1e-14
    assert np.abs(smooth).max() < 1.5 * np.abs(reference).max()
    assert np.abs(sharp).max() > 5.0 * np.abs(smooth).max()
