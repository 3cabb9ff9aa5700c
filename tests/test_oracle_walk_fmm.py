"""Pins the oracle's Barnes-Hut walk (oracle_evaluate_groups, restating
evaluate_groups_pyccel.py:16-297) and its FMM (oracle_fmm_traverse, oracle_pair_csr,
oracle_fmm_m2l_host, oracle_fmm_downward, oracle_fmm_evaluate_host, restating the
reference's Pyccel kernels of the same names).

Independent checks: theta = 0 is the exact sum; SymPy closed forms of an accepted root
cell (flat_tree.py:284-338); convergence orders of the reference's ConvergenceProblems
(flat_tree.py:341-403); the traversal covers every particle pair exactly once; L2L is an
exact polynomial re-centring; an M2L reproduces the multipole's complex potential; the
reference's accuracy tables (test_fmm.py:38-48, test_log_tree.py:19-28) and momentum
tests (test_fmm.py:85-91, :176-187).
"""

import numpy as np
import pytest
import sympy


def relative(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def cloud(n=3000, seed=0):
    rng = np.random.default_rng(seed)
    pos = np.concatenate(
        (rng.normal([[-0.3], [0.1]], 0.1, (2, n // 2)),
         rng.normal([[0.4], [-0.2]], 0.25, (2, n - n // 2))), axis=1).clip(-0.99, 0.99)
    return pos, rng.normal(0.3, 1, n)


@pytest.mark.parametrize("kernel", ["coulomb", "log"])
def test_walk_theta_zero_is_the_direct_sum(oracle, kernel):
    pos, q = cloud(500, seed=4)
    own = np.arange(500)
    t = oracle.OracleTree(pos, q, (0, 0, 1), kernel=kernel, leaf_capacity=3)
    got = t.evaluate(pos, exclude=own, theta=0.0, softening=0.01)
    want = oracle.direct_field(pos, pos, q, exclude=own, softening=0.01, kernel=kernel)
    assert relative(got[0], want[0]) < 1e-13 and relative(got[1], want[1]) < 1e-13


def multipole_rules():
    x, y, eps = sympy.symbols("x y eps", real=True)
    x1, y1, x2, y2, q1, q2 = sympy.symbols("x1 y1 x2 y2 q1 q2", real=True)
    terms = ((q1, x1, y1), (q2, x2, y2))
    charge = q1 + q2
    p_x = sum(q * a for q, a, _ in terms)
    p_y = sum(q * b for q, _, b in terms)
    s_xx = sum(q * a * a for q, a, _ in terms)
    s_xy = sum(q * a * b for q, a, b in terms)
    s_yy = sum(q * b * b for q, _, b in terms)
    rho = sympy.sqrt(x**2 + y**2 + eps**2)
    quad = s_xx * x**2 + 2 * s_xy * x * y + s_yy * y**2
    coulomb = charge / rho + (p_x * x + p_y * y) / rho**3 + (3 * quad / rho**5 - (s_xx + s_yy) / rho**3) / 2
    w = 1 / (x + sympy.I * y)
    series = sum(sum(q * (a + sympy.I * b) ** k for q, a, b in terms) * w**k / k for k in range(1, 4))
    log = -charge * sympy.log(x**2 + y**2) + 2 * sympy.re(sympy.expand_complex(series))
    args = (x, y, x1, y1, x2, y2, q1, q2, eps)
    return sympy.lambdify(args, coulomb), sympy.lambdify(args, log)


def test_walk_root_expansion_matches_sympy(oracle):
    # flat_tree.py:284-338: two charges in [-1, 1]^2 split into leaves; a target beyond
    # 2h/theta takes the root's quadrupole (coulomb) or its p = 3 series (log)
    f_coulomb, f_log = multipole_rules()
    rng = np.random.default_rng(6)
    for _ in range(20):
        x, y = rng.uniform(2.5, 4.0, 2)
        x1, y1, x2, y2 = rng.uniform(-0.9, 0.9, 4)
        q1, q2 = rng.uniform(-2, 2, 2)
        eps = rng.uniform(0, 0.5)
        pos = np.array([[x1, x2], [y1, y2]])
        for kernel, f, e in (("coulomb", f_coulomb, eps), ("log", f_log, 0.0)):
            t = oracle.OracleTree(pos, [q1, q2], (0, 0, 1), leaf_capacity=1, kernel=kernel, order=3)
            phi, _ = t.evaluate([[x], [y]], theta=1.0, softening=e)
            assert phi[0] == pytest.approx(f(x, y, x1, y1, x2, y2, q1, q2, e), rel=1e-12, abs=1e-13)


def test_walk_convergence_order(oracle):
    # flat_tree.py:341-372: error O(s^3) through the quadrupole and O(s^(p+1)) at p = 2
    rng = np.random.default_rng(1)
    offsets = rng.uniform(-0.5, 0.5, (2, 12))
    charges = rng.normal(0.3, 1.0, 12)
    angle = np.linspace(0.0, 2.0 * np.pi, 16, endpoint=False)
    targets = 4.0 * np.stack((np.cos(angle), np.sin(angle)))
    for kernel in ("coulomb", "log"):
        errors = []
        for size in (1.0, 0.5, 0.25, 0.125):
            pos = size * offsets
            t = oracle.OracleTree(pos, charges, (0, 0, 1), leaf_capacity=1, kernel=kernel, order=2)
            _, field = t.evaluate(targets, theta=0.9)
            _, exact = oracle.direct_field(targets, pos, charges, kernel=kernel)
            errors.append(np.max(np.abs(field - exact)))
        slopes = np.log2(np.array(errors[:-1]) / np.array(errors[1:]))
        assert np.all(slopes > 2.6), (kernel, slopes)


@pytest.mark.parametrize("order,tolerance", [(4, 1e-3), (8, 1e-5), (12, 2e-7)])
def test_log_walk_accuracy_table(oracle, order, tolerance):
    rng = np.random.default_rng(0)
    pos, q = rng.normal(0, 0.2, (2, 3000)).clip(-0.99, 0.99), rng.normal(0.3, 1, 3000)
    probes = np.random.default_rng(1).uniform(-1, 1, (2, 200))
    phi, field = oracle.OracleTree(pos, q, (0, 0, 1), kernel="log", order=order).evaluate(
        probes, theta=0.6)
    rp, rf = oracle.direct_field(probes, pos, q, kernel="log")
    assert relative(phi, rp) < tolerance and relative(field, rf) < tolerance


@pytest.mark.parametrize("leaf", [1, 4, 16])
def test_traversal_covers_every_pair_once(oracle, leaf):
    pos, q = cloud(300, seed=9)
    t = oracle.OracleTree(pos, q, (0, 0, 1), kernel="log", leaf_capacity=leaf)
    m2l, p2p = t.traverse(0.5)
    n = pos.shape[1]
    cover = np.zeros((n, n), dtype=np.int64)
    members = [t.perm[t.node_start[k]:t.node_end[k]] for k in range(t.node_count)]
    for a, b in m2l:
        assert a != b
        cover[np.ix_(members[a], members[b])] += 1
        cover[np.ix_(members[b], members[a])] += 1
        # well separated by the reference's criterion
        d = np.hypot(*(t.node_center[a] - t.node_center[b]))
        assert np.sqrt(2) * (t.node_half[a] + t.node_half[b]) < 0.5 * d
    for a, b in p2p:
        assert t.node_child[a] < 0 and t.node_child[b] < 0
        cover[np.ix_(members[a], members[b])] += 1
        if a != b:
            cover[np.ix_(members[b], members[a])] += 1
    # every pair i != j exactly once; (i, i) once, by its leaf's self pair
    assert np.all(cover == 1)


def test_l2l_is_an_exact_recentring(oracle):
    rng = np.random.default_rng(4)
    pos = rng.uniform(-1, 1, (2, 200))
    t = oracle.OracleTree(pos, np.ones(200), (0, 0, 1), kernel="log", order=9, leaf_capacity=4)
    p1 = 10
    lre = np.zeros((t.node_count, p1))
    lim = np.zeros((t.node_count, p1))
    lre[0] = rng.normal(size=p1)
    lim[0] = rng.normal(size=p1)
    bsz = 19
    binom = oracle.binomial_table(bsz)
    for lo, hi in zip(t.levels[:-2], t.levels[1:-1]):
        oracle.lib().oracle_fmm_downward(int(lo), int(hi), t.node_center.ctypes.data,
                                         t.node_child.ctypes.data, binom.ctypes.data, bsz, 9,
                                         lre.ctypes.data, lim.ctypes.data)
    c0 = t.node_center[0, 0] + 1j * t.node_center[0, 1]
    b0 = lre[0] + 1j * lim[0]
    for node in range(1, t.node_count):
        c = t.node_center[node, 0] + 1j * t.node_center[node, 1]
        z = c + t.node_half[node] * (rng.uniform(-1, 1) + 1j * rng.uniform(-1, 1))
        want = np.polyval(b0[::-1], z - c0)
        got = np.polyval((lre[node] + 1j * lim[node])[::-1], z - c)
        assert abs(got - want) < 1e-12 * np.abs(b0).sum() * 4


def test_m2l_reproduces_the_multipole_potential(oracle):
    # one source cell: its multipole's complex potential a_0 ln(z-c) - sum a_k/(k (z-c)^k),
    # expanded about a well separated cell, must agree to the truncation (theta^(p+1))
    rng = np.random.default_rng(5)
    order = 16
    p1 = order + 1
    cs, ct = -0.6 - 0.5j, 0.55 + 0.4j
    a = (rng.normal(size=p1) + 1j * rng.normal(size=p1)) * 0.1 ** np.arange(p1)
    centers = np.array([[ct.real, ct.imag], [cs.real, cs.imag]])
    rows = np.array([0, 1, 1], dtype=np.int64)
    sources = np.array([1], dtype=np.int64)
    mre = np.zeros((2, p1))
    mim = np.zeros((2, p1))
    mre[1], mim[1] = a.real, a.imag
    lre = np.zeros((2, p1))
    lim = np.zeros((2, p1))
    coeff = oracle.m2l_coefficients(order, False)
    oracle.lib().oracle_fmm_m2l_host(rows.ctypes.data, sources.ctypes.data, 2, centers.ctypes.data,
                                     centers.ctypes.data, mre.ctypes.data, mim.ctypes.data,
                                     coeff.ctypes.data, order, lre.ctypes.data, lim.ctypes.data)
    b = lre[0] + 1j * lim[0]
    for _ in range(10):
        z = ct + 0.1 * (rng.uniform(-1, 1) + 1j * rng.uniform(-1, 1))
        zz = z - cs
        exact = a[0] * np.log(zz) - sum(a[k] / (k * zz**k) for k in range(1, p1))
        got = np.polyval(b[::-1], z - ct)
        # Re is what the potential uses (-2 Re W); Im differs by the branch of ln
        assert abs(got.real - exact.real) < 1e-12


@pytest.mark.parametrize("order,tolerance", [(6, 1e-3), (10, 1e-5), (18, 1e-8)])
def test_fmm_accuracy_table(oracle, order, tolerance):
    pos, q = cloud()
    rp, rf = oracle.direct_field(pos, pos, q, exclude=np.arange(3000), kernel="log")
    phi, field = oracle.OracleTree(pos, q, (0, 0, 1), kernel="log", order=order).fmm(theta=0.5)
    assert relative(phi, rp) < tolerance and relative(field, rf) < 10 * tolerance


def test_fmm_convergence_order(oracle):
    # flat_tree.py:375-403: O(s^(p+1)) potential, O(s^p) field at p = 2
    rng = np.random.default_rng(2)
    centres = np.array([[-0.75, -0.25, 0.75, 0.25], [0.75, 0.25, -0.75, -0.25]])
    offsets = rng.uniform(-0.2, 0.2, (2, 4, 3))
    charges = rng.normal(0.3, 1.0, 12)
    pe, fe = [], []
    for size in (0.5, 0.25, 0.125, 0.0625):
        pos = (centres[:, :, None] + size * offsets).reshape(2, -1)
        t = oracle.OracleTree(pos, charges, (0, 0, 1), leaf_capacity=3, kernel="log", order=2)
        phi, field = t.fmm(theta=0.5)
        rp, rf = oracle.direct_field(pos, pos, charges, exclude=np.arange(12), kernel="log")
        pe.append(np.max(np.abs(phi - rp)))
        fe.append(np.max(np.abs(field - rf)))
    assert np.all(np.log2(np.array(pe[:-1]) / pe[1:]) > 2.6)
    assert np.all(np.log2(np.array(fe[:-1]) / fe[1:]) > 1.7)


def test_fmm_momentum(oracle):
    pos, q = cloud(6000, seed=3)
    _, field = oracle.OracleTree(pos, q, (0, 0, 1), kernel="log", order=18).fmm(theta=0.5)
    force = q * field
    assert np.abs(force.sum(axis=1)).max() < 1e-11 * np.abs(force).sum()
    _, sym = oracle.OracleTree(pos, q, (0, 0, 1), kernel="log", order=18, symmetric=True).fmm(theta=0.5)
    force = q * sym
    assert np.abs(force.sum(axis=1)).max() < 1e-15 * np.abs(force).sum()


def test_fmm_failures(oracle):
    pos, q = cloud(400, seed=2)
    pos[:, 2] = pos[:, 1]
    with pytest.raises(ValueError, match="Coincident"):
        oracle.OracleTree(pos, q, (0, 0, 1), kernel="log").fmm()
    oracle.OracleTree(pos, q, (0, 0, 1), kernel="log").fmm(softening=0.01)
    with pytest.raises(ValueError, match="theta"):
        oracle.OracleTree(pos, q, (0, 0, 1), kernel="log").fmm(theta=1.0)


def test_fmm_at_matches_direct_and_covers_pairs(oracle):
    # test_fmm.py:225-250 (log kernel): 2e-4 against the direct sum; the dual lists cover
    # every (target, source) pair exactly once
    rng = np.random.default_rng(13)
    pos, q = cloud(4000, seed=13)
    targets = np.concatenate((rng.uniform(-0.9, 0.9, (2, 1500)), pos[:, :500]), axis=1)
    exclude = np.concatenate((np.full(1500, -1), np.arange(500)))
    t = oracle.OracleTree(pos, q, (0, 0, 1), kernel="log", order=16, leaf_capacity=16)
    phi, field = t.fmm_at(targets, exclude=exclude, theta=0.5)
    rp, rf = oracle.direct_field(targets, pos, q, exclude=exclude, kernel="log")
    assert relative(field, rf) < 2e-4 and relative(phi, rp) < 2e-4
    small = oracle.OracleTree(pos[:, :300], q[:300], (0, 0, 1), kernel="log", leaf_capacity=4)
    tg = rng.uniform(-0.9, 0.9, (2, 200))
    small.fmm_at(tg, theta=0.5, leaf_capacity=3)
    m2l, p2p = small.dual_lists
    lo, hi = tg.min(axis=1), tg.max(axis=1)
    c = 0.5 * (lo + hi)
    tt = oracle.OracleTree(tg, np.zeros(200), (c[0], c[1], 0.5 * np.max(hi - lo) * (1 + 1e-12) + 1e-300),
                           leaf_capacity=3)
    cover = np.zeros((200, 300), dtype=np.int64)
    tm = [tt.perm[tt.node_start[k]:tt.node_end[k]] for k in range(tt.node_count)]
    sm = [small.perm[small.node_start[k]:small.node_end[k]] for k in range(small.node_count)]
    for a, b in np.concatenate((m2l, p2p)):
        cover[np.ix_(tm[a], sm[b])] += 1
    assert np.all(cover == 1)


def test_cartesian_taylor_coefficients_match_sympy(oracle):
    # the M2L's A_g(t) = D^g f(t) / g! of f = (|t|^2 + eps^2)^(-1/2) (_cartesian.py:11-14):
    # one unit moment m_0 = 1 at the source gives local T_b = A_b(t) exactly
    x, y, e = sympy.symbols("x y e", real=True)
    f = 1 / sympy.sqrt(x**2 + y**2 + e**2)
    order = 5
    tb = oracle.cart_tables(order)
    nc = oracle.cart_size(order)
    tx, ty, eps = 0.7, -0.45, 0.2
    centers = np.array([[tx, ty], [0.0, 0.0]])
    moments = np.zeros((2, nc))
    moments[1, 0] = 1.0
    local = np.zeros((2, nc))
    rows = np.array([0, 1, 1], dtype=np.int64)
    src = np.array([1], dtype=np.int64)
    oracle.lib().oracle_cart_m2l_host(
        rows.ctypes.data, src.ctypes.data, 2, centers.ctypes.data, centers.ctypes.data,
        moments.ctypes.data, nc, eps, tb["ex"].ctypes.data, tb["ey"].ctypes.data,
        tb["rec"].ctypes.data, tb["m2l"].ctypes.data, tb["m2l_coef"].ctypes.data, len(tb["m2l"]),
        local.ctypes.data)
    for i in range(nc):
        a, b = int(tb["ex"][i]), int(tb["ey"][i])
        d = sympy.diff(f, x, a, y, b) if a or b else f
        want = float(d.subs({x: tx, y: ty, e: eps})) / (sympy.factorial(a) * sympy.factorial(b))
        assert local[0, i] == pytest.approx(want, rel=1e-12, abs=1e-14)


def test_cartesian_moments_equal_brute_force_sums(oracle):
    # m_(a, b) = sum over the cell's particles of q (x - cx)^a (y - cy)^b
    # (cart_moments_upward_pyccel.py:14-82: P2M at the leaves, M2M shifts above them)
    rng = np.random.default_rng(17)
    pos, q = cloud(600, seed=17)
    order = 7
    t = oracle.OracleTree(pos, q, (0, 0, 1), order=order, leaf_capacity=5)
    moments = oracle._cart_moments(t)
    tb = oracle.cart_tables(order)
    assert int(np.max(t.node_depth)) >= 4
    internal = np.flatnonzero((t.node_child >= 0) & (t.node_end > t.node_start))
    leaves = np.flatnonzero((t.node_child < 0) & (t.node_end > t.node_start))
    nodes = np.concatenate(([0], rng.choice(internal, 25), rng.choice(leaves, 25)))
    for node in nodes:
        members = t.perm[t.node_start[node]:t.node_end[node]]
        dx = pos[0, members] - t.node_center[node, 0]
        dy = pos[1, members] - t.node_center[node, 1]
        want = np.array([np.sum(q[members] * dx ** a * dy ** b)
                         for a, b in zip(tb["ex"], tb["ey"])])
        scale = np.sum(np.abs(q[members])) * (np.sqrt(2) * t.node_half[node]) ** (tb["ex"] + tb["ey"])
        assert np.all(np.abs(moments[node] - want) <= 1e-12 * scale), node


def test_cartesian_l2l_is_an_exact_recentring(oracle):
    # cart_downward_pyccel.py:14-55: T_b(child) += sum_g C(g, b) T_g s^(g - b) re-expands
    # the parent's polynomial sum_g T_g (r - c)^g about the child's centre, exactly (the
    # degree does not grow); after the pass a cell's polynomial is the sum of its own and
    # every ancestor's starting polynomials, each about its own centre
    rng = np.random.default_rng(18)
    pos = rng.uniform(-1, 1, (2, 300))
    order = 6
    t = oracle.OracleTree(pos, np.ones(300), (0, 0, 1), order=order, leaf_capacity=4)
    tb = oracle.cart_tables(order)
    nc = oracle.cart_size(order)
    start = rng.normal(size=(t.node_count, nc))
    local = start.copy()
    lev = np.ascontiguousarray(t.levels, dtype=np.int64)
    oracle.lib().oracle_cart_downward(
        lev.ctypes.data, len(lev) - 1, t.node_center.ctypes.data, t.node_child.ctypes.data,
        tb["mono_parent"].ctypes.data, tb["mono_dir"].ctypes.data, tb["shift"].ctypes.data,
        tb["shift_coef"].ctypes.data, len(tb["shift"]), nc, local.ctypes.data)
    parent = np.full(t.node_count, -1)
    for node in range(t.node_count):
        if t.node_child[node] >= 0:
            parent[t.node_child[node]:t.node_child[node] + 4] = node

    def poly(coef, centre, r):
        d = r - centre
        return np.sum(coef * d[0] ** tb["ex"] * d[1] ** tb["ey"])

    for node in rng.choice(t.node_count, 60):
        r = t.node_center[node] + t.node_half[node] * rng.uniform(-1, 1, 2)
        want, a = 0.0, node
        while a >= 0:
            want += poly(start[a], t.node_center[a], r)
            a = parent[a]
        got = poly(local[node], t.node_center[node], r)
        assert abs(got - want) < 1e-12 * (t.node_depth[node] + 1) * nc, node


@pytest.mark.parametrize("softening", [0.0, 0.02])
def test_coulomb_fmm_converges_and_conserves_momentum(oracle, softening):
    # test_fmm.py:189-207
    rng = np.random.default_rng(11)
    pos = np.concatenate((rng.uniform(-0.9, 0.9, (2, 1500)), rng.normal(0.3, 0.05, (2, 1500))),
                         axis=1).clip(-0.99, 0.99)
    q = rng.normal(size=pos.shape[1])
    own = np.arange(pos.shape[1])
    pe, exact = oracle.direct_field(pos, pos, q, exclude=own, softening=softening)
    errors = []
    for order in (2, 4, 6, 8):
        t = oracle.OracleTree(pos, q, (0, 0, 1), order=order, leaf_capacity=16)
        phi, field = t.fmm(theta=0.5, softening=softening)
        errors.append(relative(field, exact))
        assert np.max(np.abs(phi - pe)) < 0.05 * np.max(np.abs(pe))
        net = np.abs((q * field).sum(axis=1)).max()
        assert net < 1e-12 * np.abs(q * field).sum()
    assert errors[-1] < 1e-2 * errors[0] and np.all(np.diff(errors) < 0)


def test_coulomb_fmm_at(oracle):
    rng = np.random.default_rng(13)
    pos, q = cloud(4000, seed=13)
    targets = np.concatenate((rng.uniform(-0.9, 0.9, (2, 1500)), pos[:, :500]), axis=1)
    exclude = np.concatenate((np.full(1500, -1), np.arange(500)))
    t = oracle.OracleTree(pos, q, (0, 0, 1), order=8, leaf_capacity=16)
    phi, field = t.fmm_at(targets, exclude=exclude, theta=0.5, softening=0.01)
    rp, rf = oracle.direct_field(targets, pos, q, exclude=exclude, softening=0.01)
    assert relative(field, rf) < 2e-4 and relative(phi, rp) < 2e-4
