"""Test infrastructure (the oracle): the smooth-membership line-charge FMM of the reference,
smooth.py:1-476, restated in FP64 NumPy for the tests of the GPU's version.  Nothing in the
package imports this file.

The interface is the reference's (BoxTree, SmoothFMM with .energy / .fields() /
.potential_at, direct_energy, direct_fields, comoving_fields) so that the reference's own
tests (tests/test_smooth.py) read the same against it.  The code is organised differently:
cells come from a breadth-first split, memberships from a per-level frontier, the lists
from a recursive dual walk, and the near field is summed by "pull" over each membership's
partners (the same routes as the reference's pairwise push, smooth.py:356-379).

Pinned (tests/test_oracle_smooth.py): the partition of unity and the band limit
(test_smooth.py:40-49), band 0 against the direct sum (:52-60), forces against central
differences of the energy (:63-86), exact antisymmetry of the triangle truncation and the
accuracy ordering of the two truncations (:89-104), accuracy with smooth membership
(:107-119), the compact softening against its closed form from Gauss's law
(smooth.py:204-221), the membership weights against a brute-force scan over every leaf,
and the lists against a brute-force check of every pair of occupied cells.
"""

from __future__ import annotations

from math import comb, sqrt

import numpy as np

ROOT2 = sqrt(2.0)
MAX_BOX_DEPTH = 40  # smooth.py:97: cells split while deeper than this is not reached


# ------------------------------------------------------------------------------ kernels
def compact_log(dz, eps2):
    """(Phi, dPhi/dx + i dPhi/dy) of the pair kernel of a compact cloud of radius
    sqrt(eps2) with density (1 - r^2/a^2)^2 (smooth.py:52-72): ln r outside the cloud and
    ln a - 11/12 + (3u - 1.5u^2 + u^3/3)/2 inside, u = r^2/a^2; eps2 = 0 is the bare ln r."""
    dz = np.asarray(dz, dtype=complex)
    r2 = dz.real * dz.real + dz.imag * dz.imag
    if eps2 <= 0.0:
        return 0.5 * np.log(r2), dz / r2
    phi = np.empty(r2.shape)
    grad = np.empty(r2.shape, dtype=complex)
    cloud = r2 < eps2
    out = ~cloud
    phi[out] = 0.5 * np.log(r2[out])
    grad[out] = dz[out] / r2[out]
    u = r2[cloud] / eps2
    phi[cloud] = 0.5 * np.log(eps2) - 11.0 / 12.0 + (3.0 * u - 1.5 * u * u + u ** 3 / 3.0) / 2.0
    grad[cloud] = dz[cloud] * (3.0 - 3.0 * u + u * u) / eps2
    return phi, grad


def quintic_step(t):
    """s(t) = 6u^5 - 15u^4 + 10u^3 with u = (t + 1)/2 clipped to [0, 1], and ds/dt
    (smooth.py:75-80); s(t) + s(-t) = 1."""
    u = np.minimum(np.maximum((np.asarray(t, dtype=float) + 1.0) / 2.0, 0.0), 1.0)
    s = u * u * u * (10.0 + u * (-15.0 + 6.0 * u))
    ds_du = 30.0 * u * u * (1.0 - u) ** 2
    return s, ds_du / 2.0


# ------------------------------------------------------------------------------ cells
class BoxTree:
    """Static square cells (smooth.py:83-121): a cell with more than leaf_capacity of the
    build points splits into (-x,+y), (+x,+y), (-x,-y), (+x,-y) children, while its depth is
    below 40; a point goes right when x > cx and up when y > cy.  Numbered breadth first."""

    def __init__(self, pos, root, *, leaf_capacity=16):
        pos = np.asarray(pos, dtype=float)
        xs, ys = pos[0], pos[1]
        cx, cy, half, parent, first_child, depth = [root.xmid], [root.ymid], [root.half], [-1], [-1], [0]
        points = [np.arange(xs.shape[0])]
        k = 0
        while k < len(cx):  # breadth first: children are appended behind every older cell
            inside = points[k]
            if inside.shape[0] > leaf_capacity and depth[k] < MAX_BOX_DEPTH:
                first_child[k] = len(cx)
                h = 0.5 * half[k]
                right = xs[inside] > cx[k]
                up = ys[inside] > cy[k]
                for go_right, go_up in ((False, True), (True, True), (False, False), (True, False)):
                    cx.append(cx[k] + (h if go_right else -h))
                    cy.append(cy[k] + (h if go_up else -h))
                    half.append(h)
                    parent.append(k)
                    first_child.append(-1)
                    depth.append(depth[k] + 1)
                    points.append(inside[(right == go_right) & (up == go_up)])
            k += 1
        self.center = np.array(cx) + 1j * np.array(cy)
        self.half = np.array(half)
        self.parent = np.array(parent, dtype=np.int64)
        self.child = np.array(first_child, dtype=np.int64)
        self.depth = np.array(depth, dtype=np.int64)
        self.root = root
        self.leaves = np.flatnonzero(self.child < 0)
        self.min_leaf_half = float(self.half[self.leaves].min())

    def hard_leaf(self, z):
        """The one leaf of each point at band 0: right when x > cx, down when y <= cy
        (smooth.py:125-143)."""
        at = np.zeros(len(z), dtype=np.int64)
        while True:
            inner = self.child[at] >= 0
            if not inner.any():
                return at
            c = self.center[at[inner]]
            quad = (z[inner].real > c.real).astype(np.int64) + 2 * (z[inner].imag <= c.imag)
            at[inner] = self.child[at[inner]] + quad

    def memberships(self, z, band):
        """(particle, leaf, w, dw/dx, dw/dy) of every particle-leaf pair with w > 0
        (smooth.py:122-193)."""
        z = np.asarray(z, dtype=complex)
        n = len(z)
        if band <= 0.0:
            ones, zeros = np.ones(n), np.zeros(n)
            return np.arange(n), self.hard_leaf(z), ones, zeros, zeros
        if band > self.min_leaf_half:
            raise ValueError("band must not exceed the smallest leaf half-width")
        found_p, found_l = [], []
        who, at = np.arange(n), np.zeros(n, dtype=np.int64)
        while who.shape[0]:
            kids = self.child[at]
            done = kids < 0
            found_p.append(who[done])
            found_l.append(at[done])
            who, kids = who[~done], kids[~done]
            keep_p, keep_n = [], []
            for q in range(4):
                cell = kids + q
                reach = self.half[cell] + band
                c = self.center[cell]
                ok = (np.abs(z[who].real - c.real) < reach) & (np.abs(z[who].imag - c.imag) < reach)
                keep_p.append(who[ok])
                keep_n.append(cell[ok])
            who, at = np.concatenate(keep_p), np.concatenate(keep_n)
        part = np.concatenate(found_p)
        leaf = np.concatenate(found_l)
        c, h, zz = self.center[leaf], self.half[leaf], z[part]
        r = self.root
        # signed distance into the leaf across its four edges, in bands, and whether the
        # edge lies on the root's boundary
        edges = [
            ((zz.real - (c.real - h)) / band, c.real - h <= r.xmid - r.half + 1e-12 * r.half),
            (((c.real + h) - zz.real) / band, c.real + h >= r.xmid + r.half - 1e-12 * r.half),
            ((zz.imag - (c.imag - h)) / band, c.imag - h <= r.ymid - r.half + 1e-12 * r.half),
            (((c.imag + h) - zz.imag) / band, c.imag + h >= r.ymid + r.half - 1e-12 * r.half),
        ]
        steps = []
        for t, on_root_edge in edges:
            s, ds = quintic_step(t)
            # an edge on the root's boundary has no neighbour to blend with: weight 1
            steps.append((np.where(on_root_edge, 1.0, s), np.where(on_root_edge, 0.0, ds)))
        (s_l, d_l), (s_r, d_r), (s_b, d_b), (s_t, d_t) = steps
        wx, wy = s_l * s_r, s_b * s_t
        w = wx * wy
        dwx = (d_l * s_r - s_l * d_r) / band * wy
        dwy = (d_b * s_t - s_b * d_t) / band * wx
        keep = w > 0.0
        return part[keep], leaf[keep], w[keep], dwx[keep], dwy[keep]


# ------------------------------------------------------------------------------ direct sums
def direct_fields(pos, charge, softening=0.0):
    """Potential (N,) and field (2, N) at every particle from all the others
    (smooth.py:431-440): phi_i = -2 sum_j q_j Phi_ij, E_i = 2 sum_j q_j grad Phi_ij."""
    pos = np.asarray(pos, dtype=float)
    charge = np.asarray(charge, dtype=float)
    z = pos[0] + 1j * pos[1]
    n = len(z)
    phi, field = np.zeros(n), np.zeros(n, dtype=complex)
    for i in range(n):
        others = np.arange(n) != i
        p, g = compact_log(z[i] - z[others], softening ** 2)
        phi[i] = -2.0 * np.dot(p, charge[others])
        field[i] = 2.0 * np.dot(g, charge[others])
    return phi, np.stack((field.real, field.imag))


def direct_energy(pos, charge, softening=0.0):
    """-2 sum_{i<j} q_i q_j Phi(z_i - z_j) (smooth.py:196-201)."""
    pos = np.asarray(pos, dtype=float)
    charge = np.asarray(charge, dtype=float)
    z = pos[0] + 1j * pos[1]
    total = 0.0
    for i in range(len(z) - 1):
        p, _ = compact_log(z[i] - z[i + 1:], softening ** 2)
        total += np.dot(p, charge[i + 1:]) * charge[i]
    return float(-2.0 * total)


def softening_closed_form(r, a):
    """Gauss's law for the compact cloud (smooth.py:204-221): ln a minus the integral of the
    enclosed fraction 1 - (1 - s^2/a^2)^3 over ds/s from r to a, for r < a, done by hand:
    (1 - (1 - x)^3)/s with x = s^2/a^2 integrates to (3x - 3x^2/2 + x^3/3)/2."""
    def primitive(s):
        x = s * s / (a * a)
        return (3.0 * x - 1.5 * x * x + x ** 3 / 3.0) / 2.0
    return np.log(a) - (primitive(a) - primitive(r))


# ------------------------------------------------------------------------------ the FMM
def dual_lists(tree, occupied, theta):
    """M2L and P2P cell pairs of the dual walk over occupied cells (smooth.py:246-277):
    a leaf with itself is near; a cell with itself opens into its ten child pairs; two
    cells are far when sqrt(2) (h_a + h_b) < theta |c_a - c_b|, near when both are leaves,
    otherwise the larger (a on ties), or the one that is not a leaf, opens."""
    far, near = [], []

    def visit(a, b):
        if not (occupied[a] and occupied[b]):
            return
        la, lb = tree.child[a] < 0, tree.child[b] < 0
        if a == b:
            if la:
                near.append((a, a))
                return
            f = tree.child[a]
            for i in range(4):
                for j in range(i, 4):
                    visit(f + i, f + j)
            return
        if ROOT2 * (tree.half[a] + tree.half[b]) < theta * abs(tree.center[a] - tree.center[b]):
            far.append((a, b))
        elif la and lb:
            near.append((a, b))
        elif lb or (not la and tree.half[a] >= tree.half[b]):
            for i in range(4):
                visit(tree.child[a] + i, b)
        else:
            for i in range(4):
                visit(a, tree.child[b] + i)

    visit(0, 0)
    return (np.array(far, dtype=np.int64).reshape(-1, 2),
            np.array(near, dtype=np.int64).reshape(-1, 2))


def _m2l_matrix(p, symmetric):
    """M[l-1, k-1] = (-1)^l C(k+l-1, l) / k for l, k = 1..p, zero where k + l > p when the
    truncation is the triangle (smooth.py:309-323)."""
    m = np.zeros((p, p))
    for ell in range(1, p + 1):
        for k in range(1, p + 1):
            if not symmetric or k + ell <= p:
                m[ell - 1, k - 1] = (-1.0) ** ell * comb(k + ell - 1, ell) / k
    return m


class SmoothFMM:
    """Energy and exact-gradient fields of line charges with smooth cell membership on a
    BoxTree (smooth.py:242-428)."""

    def __init__(self, tree, pos, charge, *, order=12, theta=0.5, band=0.0, symmetric=True,
                 softening=0.0):
        self.tree = tree
        self.z = np.asarray(pos[0], dtype=float) + 1j * np.asarray(pos[1], dtype=float)
        self.q = np.asarray(charge, dtype=float)
        self.order = int(order)
        self.theta = float(theta)
        self.band = float(band)
        self.symmetric = bool(symmetric)
        self.eps2 = float(softening) ** 2
        self._solve()

    def _solve(self):
        t, p, z = self.tree, self.order, self.z
        part, leaf, w, dwx, dwy = t.memberships(z, self.band)
        self.part, self.leaf, self.w, self.dw = part, leaf, w, dwx + 1j * dwy
        qw = self.q[part] * w
        cells = len(t.half)
        # moments a_k = sum q w (z - c)^k over the memberships below each cell
        moments = np.zeros((cells, p + 1), dtype=complex)
        occupied = np.zeros(cells, dtype=bool)
        for m in range(len(part)):
            cell = leaf[m]
            while cell >= 0:
                moments[cell] += qw[m] * (z[part[m]] - t.center[cell]) ** np.arange(p + 1)
                occupied[cell] = True
                cell = t.parent[cell]
        self.moments = moments
        far, near = dual_lists(t, occupied, self.theta)
        self.m2l_pairs, self.p2p_pairs = far, near
        # M2L both ways: b_0 += a_0 ln t - sum_k a_k / (k t^k);
        # b_l += ((-1)^(l+1) a_0 / l - sum_k M[l, k] a_k / t^k) / t^l
        local = np.zeros((cells, p + 1), dtype=complex)
        mat = _m2l_matrix(p, self.symmetric)
        ks = np.arange(1, p + 1)
        for src, dst in list(far) + [(b, a) for a, b in far]:
            tt = t.center[dst] - t.center[src]
            inv_k = (1.0 / tt) ** ks
            a = moments[src]
            local[dst, 0] += a[0] * np.log(tt) - np.sum(a[1:] * inv_k / ks)
            local[dst, 1:] += ((-1.0) ** (ks + 1) * a[0] / ks - mat @ (a[1:] * inv_k)) * inv_k
        # L2L: b'_m = sum_{l >= m} C(l, m) b_l s^(l - m), s = c_child - c_parent, from the root
        for cell in np.argsort(t.depth, kind="stable"):
            f = t.child[cell]
            if f < 0:
                continue
            for c in range(f, f + 4):
                s = t.center[c] - t.center[cell]
                shifted = np.zeros(p + 1, dtype=complex)
                for ell in range(p + 1):
                    for mm in range(ell + 1):
                        shifted[mm] += comb(ell, mm) * local[cell, ell] * s ** (ell - mm)
                local[c] += shifted
        self.local = local
        # Phi_m, G_m: the leaf's local expansion at the particle (Horner for W and W') ...
        n_m = len(part)
        phi_m = np.zeros(n_m)
        g_m = np.zeros(n_m, dtype=complex)
        members_of = {}
        for m in range(n_m):
            members_of.setdefault(int(leaf[m]), []).append(m)
        partners = {}
        for a, b in near:
            partners.setdefault(int(a), set()).add(int(b))
            partners.setdefault(int(b), set()).add(int(a))
        for m in range(n_m):
            u = z[part[m]] - t.center[leaf[m]]
            wv, dv = 0.0 + 0.0j, 0.0 + 0.0j
            for ell in range(p, -1, -1):
                dv = dv * u + wv
                wv = wv * u + local[leaf[m], ell]
            # ... plus the exact kernel over the memberships of the near leaves (other
            # particles only; a blended particle never meets itself)
            others = [j for b in partners.get(int(leaf[m]), ()) for j in members_of.get(b, ())
                      if part[j] != part[m]]
            acc_p, acc_g = 0.0, 0.0 + 0.0j
            if others:
                others = np.array(others, dtype=np.int64)
                kp, kg = compact_log(z[part[m]] - z[part[others]], self.eps2)
                acc_p = np.dot(kp, qw[others])
                acc_g = np.dot(kg, qw[others])
            phi_m[m] = wv.real + acc_p
            g_m[m] = np.conj(dv) + acc_g
        self.phi_m, self.g_m = phi_m, g_m

    @property
    def energy(self):
        """U = -sum_m q w Phi_m (smooth.py:410-413)."""
        return float(-np.sum(self.q[self.part] * self.w * self.phi_m))

    def fields(self):
        """phi_i = -2 sum_m w Phi_m, E_i = 2 sum_m (w G_m + Phi_m grad w) (smooth.py:415-422)."""
        n = len(self.z)
        phi = np.bincount(self.part, -2.0 * self.w * self.phi_m, minlength=n)
        e = 2.0 * (self.w * self.g_m + self.phi_m * self.dw)
        ex = np.bincount(self.part, e.real, minlength=n)
        ey = np.bincount(self.part, e.imag, minlength=n)
        return phi, np.stack((ex, ey))

    def potential_at(self, points):
        """-2 sum_j q_j Phi(z - z_j) at external points (smooth.py:424-428)."""
        zp = np.asarray(points[0], dtype=float) + 1j * np.asarray(points[1], dtype=float)
        out = np.empty(len(zp))
        for i, zi in enumerate(zp):
            out[i] = -2.0 * np.dot(compact_log(zi - self.z, self.eps2)[0], self.q)
        return out


def comoving_fields(tree, pos, charge, mass, center, **options):
    """Energy, potential and field with the cells following the centre of mass
    (smooth.py:453-476): the tree solves at x - (c(x) - center), and the field loses each
    particle's mass share of the net force, divided by its charge."""
    pos = np.asarray(pos, dtype=float)
    mass = np.asarray(mass, dtype=float)
    charge = np.asarray(charge, dtype=float)
    com = (pos * mass).sum(axis=1) / mass.sum()
    shift = com - np.asarray(center, dtype=float)
    f = SmoothFMM(tree, pos - shift[:, None], charge, **options)
    phi, field = f.fields()
    net = (field * charge).sum(axis=1)
    field = field - np.outer(net, mass / mass.sum()) / charge
    return f.energy, phi, field
