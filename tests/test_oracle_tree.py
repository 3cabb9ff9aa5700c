"""Pins the oracle's tree build (oracle_quadtree_build, restating
quadtree_build_pyccel.py:12-137) and its moments (log_moments_upward_pyccel.py:14-93,
quadtree_moments_upward_pyccel.py:13-83).

The tree is checked against a recursive depth-first construction written here from the
rules alone (a different algorithm from the oracle's breadth-first stable partition),
and against the invariants that fix the arrays uniquely.  Moments are checked against a
brute-force sum over each cell's particles (the oracle shifts children up instead).
"""

import numpy as np
import pytest
import sympy


def recursive_tree(pos, root, cap, max_depth):
    """Nodes as (centre, half, depth, particle list); children created for every split."""
    nodes = []

    def build(idx, cx, cy, h, depth, out):
        out.append([cx, cy, h, depth, idx, None])
        me = out[-1]
        if len(idx) > cap and depth < max_depth:
            hc = 0.5 * h
            groups = [[], [], [], []]
            for j in idx:
                q = (1 if pos[0, j] > cx else 0) + (2 if pos[1, j] < cy else 0)
                groups[q].append(j)
            me[5] = [(cx - hc, cy + hc), (cx + hc, cy + hc), (cx - hc, cy - hc), (cx + hc, cy - hc)]
            me.append(groups)
        return me

    # breadth-first numbering of a depth-first description
    top = build(list(range(pos.shape[1])), root[0], root[1], root[2], 0, nodes)
    order = [top]
    i = 0
    while i < len(order):
        node = order[i]
        if node[5] is not None:
            for (cx, cy), grp in zip(node[5], node[6]):
                child = []
                build(grp, cx, cy, 0.5 * node[2], node[3] + 1, child)
                order.append(child[0])
            node.append(len(order) - 4)
        i += 1
    return order


def flatten(order):
    perm = []
    start = {}
    # tree order: a node's particles are the concatenation of its children's
    def collect(k):
        node = order[k]
        if node[5] is None:
            return list(node[4])
        first = node[-1]
        out = []
        for c in range(4):
            out += collect(first + c)
        return out

    perm = collect(0)
    return perm


CLOUDS = {
    "uniform": lambda r: r.uniform(-0.99, 0.99, (2, 300)),
    "clustered": lambda r: np.clip(r.normal(0.3, 0.01, (2, 200)), -1, 1),
    "lattice": lambda r: np.stack(np.meshgrid(np.linspace(-0.5, 0.5, 9),
                                              np.linspace(-0.5, 0.5, 9))).reshape(2, -1),
    "coincident": lambda r: np.full((2, 12), 0.25),
}


@pytest.mark.parametrize("name", list(CLOUDS))
@pytest.mark.parametrize("cap,max_depth", [(1, 10), (6, 48), (3, 5)])
def test_tree_equals_recursive_construction(oracle, name, cap, max_depth):
    pos = CLOUDS[name](np.random.default_rng(5))
    tree = oracle.OracleTree(pos, np.ones(pos.shape[1]), (0.0, 0.0, 1.0), leaf_capacity=cap,
                             max_depth=max_depth)
    order = recursive_tree(pos, (0.0, 0.0, 1.0), cap, max_depth)
    assert tree.node_count == len(order)
    np.testing.assert_array_equal(tree.node_center, [[n[0], n[1]] for n in order])
    np.testing.assert_array_equal(tree.node_half, [n[2] for n in order])
    np.testing.assert_array_equal(tree.node_depth, [n[3] for n in order])
    np.testing.assert_array_equal(tree.node_child, [n[-1] if n[5] is not None else -1 for n in order])
    np.testing.assert_array_equal(tree.perm, flatten(order))
    sizes = tree.node_end - tree.node_start
    np.testing.assert_array_equal(sizes, [len(n[4]) for n in order])
    # levels bound the depths
    for d in range(len(tree.levels) - 1):
        assert np.all(tree.node_depth[tree.levels[d]:tree.levels[d + 1]] == d)


def test_tree_invariants_on_a_larger_cloud(oracle):
    rng = np.random.default_rng(7)
    pos = np.concatenate((rng.uniform(-1, 1, (2, 3000)), rng.normal(0.2, 0.01, (2, 2000))),
                         axis=1).clip(-1, 1)
    cap = 8
    t = oracle.OracleTree(pos, np.ones(pos.shape[1]), (0.0, 0.0, 1.0), leaf_capacity=cap)
    assert np.array_equal(np.sort(t.perm), np.arange(pos.shape[1]))
    for node in range(t.node_count):
        s, e, c = t.node_start[node], t.node_end[node], t.node_child[node]
        if c < 0:
            assert e - s <= cap or t.node_depth[node] == 48
            assert np.all(np.diff(t.perm[s:e]) > 0)  # input order inside a leaf
            continue
        assert e - s > cap
        assert t.node_start[c] == s and t.node_end[c + 3] == e
        for q in range(4):
            k = c + q
            if q:
                assert t.node_start[k] == t.node_end[k - 1]
            assert t.node_half[k] == 0.5 * t.node_half[node]
            sx = 1.0 if q & 1 else -1.0
            sy = -1.0 if q & 2 else 1.0
            assert t.node_center[k, 0] == t.node_center[node, 0] + sx * t.node_half[k]
            assert t.node_center[k, 1] == t.node_center[node, 1] + sy * t.node_half[k]
            members = t.perm[t.node_start[k]:t.node_end[k]]
            assert np.all((pos[0, members] > t.node_center[node, 0]) == bool(q & 1))
            assert np.all((pos[1, members] < t.node_center[node, 1]) == bool(q & 2))


def test_coincident_chain_and_capacity_restart(oracle):
    t = oracle.OracleTree([[0.3, 0.3, -0.5], [0.3, 0.3, 0.2]], [1, 1, -1], (0, 0, 1),
                          leaf_capacity=1, max_depth=48)
    assert t.node_count == 1 + 4 * 48 and t.node_depth.max() == 48
    rng = np.random.default_rng(5)
    pos = np.clip(rng.normal(0.3, 0.01, (2, 2000)), -1, 1)
    big = oracle.OracleTree(pos, np.ones(2000), (0, 0, 1), leaf_capacity=6)
    small = oracle.OracleTree.__new__(oracle.OracleTree)
    small.__dict__.update(big.__dict__)
    small.pos = big.pos
    small._build(0.0, 0.0, 1.0, 8)  # a too-small node buffer restarts the build
    assert small.build_calls > 1
    np.testing.assert_array_equal(small.node_child, big.node_child)
    np.testing.assert_array_equal(small.perm, big.perm)


@pytest.mark.parametrize("kernel", ["coulomb", "log"])
def test_moments_equal_brute_force_per_cell(oracle, kernel):
    rng = np.random.default_rng(3)
    pos = np.concatenate((rng.uniform(-1, 1, (2, 800)), rng.normal(-0.4, 0.03, (2, 400))),
                         axis=1).clip(-1, 1)
    q = rng.normal(0.2, 1.0, pos.shape[1])
    q[::5] = 0.0
    t = oracle.OracleTree(pos, q, (0, 0, 1), kernel=kernel, order=10, leaf_capacity=5)
    z = pos[0] + 1j * pos[1]
    for node in range(t.node_count):
        members = t.perm[t.node_start[node]:t.node_end[node]]
        c = t.node_center[node, 0] + 1j * t.node_center[node, 1]
        d = z[members] - c
        qm = q[members]
        scale = np.abs(qm).sum() + 1e-300
        assert t.node_qabs[node] == pytest.approx(np.abs(qm).sum(), rel=1e-13, abs=1e-13)
        if kernel == "log":
            a = np.array([(qm * d**k).sum() for k in range(11)])
            got = t.moment_re[node] + 1j * t.moment_im[node]
            h = np.sqrt(2) * t.node_half[node]
            assert np.max(np.abs((got - a) / h ** np.arange(11))) < 1e-12 * scale
        else:
            dx, dy = d.real, d.imag
            want = [qm.sum(), (qm * dx).sum(), (qm * dy).sum(), (qm * dx * dx).sum(),
                    (qm * dx * dy).sum(), (qm * dy * dy).sum()]
            got = [t.node_q[node], *t.node_p[node], *t.node_s[node]]
            np.testing.assert_allclose(got, want, rtol=0, atol=1e-13 * scale)


def test_moment_rules_sympy(oracle):
    # flat_tree.py:248-281: two charges in a root cell, leaf_capacity 1, order 3
    cx, cy, x1, y1, x2, y2, q1, q2 = sympy.symbols("cx cy x1 y1 x2 y2 q1 q2", real=True)
    terms = ((q1, x1, y1), (q2, x2, y2))
    s_xy = sum(q * (x - cx) * (y - cy) for q, x, y in terms)
    a_3 = sympy.im(sympy.expand(sum(q * (x - cx + sympy.I * (y - cy)) ** 3 for q, x, y in terms)))
    f_s = sympy.lambdify((cx, cy, x1, y1, x2, y2, q1, q2), s_xy)
    f_a = sympy.lambdify((cx, cy, x1, y1, x2, y2, q1, q2), a_3)
    rng = np.random.default_rng(2)
    for _ in range(20):
        args = (*rng.uniform(-0.1, 0.1, 2), *rng.uniform(-0.85, 0.85, 4), *rng.uniform(-2, 2, 2))
        cxv, cyv, a1, b1, a2, b2, c1, c2 = args
        pos = np.array([[a1, a2], [b1, b2]])
        tc = oracle.OracleTree(pos, [c1, c2], (cxv, cyv, 1.0), leaf_capacity=1, order=3)
        tl = oracle.OracleTree(pos, [c1, c2], (cxv, cyv, 1.0), leaf_capacity=1, order=3,
                               kernel="log")
        assert tc.node_s[0, 1] == pytest.approx(f_s(*args), rel=1e-12, abs=1e-14)
        assert tl.moment_im[0, 3] == pytest.approx(f_a(*args), rel=1e-12, abs=1e-14)
