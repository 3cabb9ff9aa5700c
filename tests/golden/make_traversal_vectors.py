"""Writes traversal.npz: the reference's own tree arrays and FMM interaction lists on small
inputs, as golden vectors for the oracle and the GPU.

Only the three kernels these vectors come from are run, each loaded by file path (they
import nothing but ``math``): ``quadtree_build`` (kernels/quadtree_build/
quadtree_build_pyccel.py:12-137), ``fmm_traverse`` (kernels/fmm_traverse/
fmm_traverse_pyccel.py:15-106) and ``dual_traverse`` (kernels/dual_traverse/
dual_traverse_pyccel.py:14-86), called as the reference's orchestration calls them:
``FlatQuadtree._build_host`` (flat_tree.py:483-530), ``_traverse`` (flat_tree.py:
1318-1358) generation by generation from (0, 0), and ``_dual_lists`` / ``_target_plan``
(flat_tree.py:1086-1149) with the target tree's non-dyadic root.  The mathematics leaves
the exact pair sets open (any cover of all particle pairs by separated cells would do), so
only the reference itself pins them.

Run in a container that holds the reference at /root/reference (never on a machine
without it; the tests read only the .npz):
    python tests/golden/make_traversal_vectors.py
"""

from __future__ import annotations

import importlib.util
import os

import numpy as np

REF = "/root/reference/src/tree_d_3d/kernels"
HERE = os.path.dirname(os.path.abspath(__file__))


def _kernel(folder: str, name: str):
    path = os.path.join(REF, folder, f"{folder}_pyccel.py")
    spec = importlib.util.spec_from_file_location(f"_ref_{folder}", path)
    module = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(module)
    return getattr(module, name)


quadtree_build = _kernel("quadtree_build", "quadtree_build")
fmm_traverse = _kernel("fmm_traverse", "fmm_traverse")
dual_traverse = _kernel("dual_traverse", "dual_traverse")


def build(pos, root, leaf, max_depth=48):
    """flat_tree.py:483-530 (_build_host) with capacity 4 (2N/leaf + 16), doubled on a
    full node array."""
    n = pos.shape[1]
    capacity = 4 * (2 * n // leaf + 16)
    while True:
        sorted_pos = np.zeros((2, n + 1))
        sorted_pos[:, :n] = pos
        center = np.empty((capacity, 2))
        half = np.empty(capacity)
        start, end, child, depth = (np.empty(capacity, dtype=np.int64) for _ in range(4))
        center[0] = root[:2]
        half[0] = root[2]
        perm, slot, quadrant, new_perm, new_slot = (np.empty(n, dtype=np.int64)
                                                    for _ in range(5))
        levels = np.empty(capacity + 1, dtype=np.int64)
        sizes = np.zeros(3, dtype=np.int64)
        quadtree_build(sorted_pos, leaf, max_depth, center, half, start, end, child, depth, perm,
                       levels, slot, quadrant, np.empty(capacity, dtype=np.int64),
                       np.empty((2, n + 1)), new_perm, new_slot, sizes)
        if not sizes[2]:
            break
        capacity *= 2
    c = int(sizes[0])
    return dict(node_center=center[:c].copy(), node_half=half[:c].copy(),
                node_start=start[:c].copy(), node_end=end[:c].copy(), node_child=child[:c].copy(),
                node_depth=depth[:c].copy(), perm=perm, levels=levels[: sizes[1] + 1].copy())


def traverse(t, theta):
    """flat_tree.py:1318-1358: generations of fmm_traverse from (0, 0), regrown on a full
    buffer."""
    c = len(t["node_half"])
    capacity = np.array([16, 16, 10]) * c + 64
    while True:
        m2l = np.empty((capacity[0], 2), dtype=np.int64)
        p2p = np.empty((capacity[1], 2), dtype=np.int64)
        frontier = np.zeros((capacity[2], 2), dtype=np.int64)
        nxt = np.empty((capacity[2], 2), dtype=np.int64)
        counts = np.zeros(4, dtype=np.int64)
        n_frontier = 1
        while n_frontier:
            counts[2] = 0
            fmm_traverse(t["node_center"], t["node_half"], t["node_start"], t["node_end"],
                         t["node_child"], theta, frontier, n_frontier, nxt, m2l, p2p, counts)
            if counts[3]:
                break
            n_frontier = int(counts[2])
            frontier, nxt = nxt, frontier
        if not counts[3]:
            return m2l[: counts[0]].copy(), p2p[: counts[1]].copy()
        capacity = np.maximum(2 * capacity, counts[:3])


def dual(t, s, theta):
    """flat_tree.py:1086-1122 (_dual_lists)."""
    depth = int(t["node_depth"].max()) + int(s["node_depth"].max())
    stack = np.empty((3 * depth + 8, 2), dtype=np.int64)
    capacity = np.array([16, 8]) * (len(t["node_half"]) + len(s["node_half"])) + 64
    counts = np.zeros(3, dtype=np.int64)
    while True:
        m2l = np.empty((capacity[0], 2), dtype=np.int64)
        p2p = np.empty((capacity[1], 2), dtype=np.int64)
        dual_traverse(t["node_center"], t["node_half"], t["node_start"], t["node_end"],
                      t["node_child"], s["node_center"], s["node_half"], s["node_start"],
                      s["node_end"], s["node_child"], theta, stack, m2l, p2p, counts)
        if not counts[2]:
            return m2l[: counts[0]].copy(), p2p[: counts[1]].copy()
        capacity = np.maximum(2 * capacity, counts[:2] + 64)


def target_root(targets):
    """flat_tree.py:1135-1137: the target tree's root square."""
    lo, hi = targets.min(axis=1), targets.max(axis=1)
    centre = 0.5 * (lo + hi)
    half = 0.5 * float(np.max(hi - lo)) * (1 + 1e-12) + 1e-300
    return np.array([centre[0], centre[1], half])


def cloud(n, seed):
    """Two Gaussian blobs (the reference's test_fmm.py:17-27 shape)."""
    rng = np.random.default_rng(seed)
    pos = np.concatenate(
        (rng.normal([[-0.3], [0.1]], 0.1, (2, n // 2)),
         rng.normal([[0.4], [-0.2]], 0.25, (2, n - n // 2))), axis=1).clip(-0.99, 0.99)
    return np.ascontiguousarray(pos)


def clustered(n, seed):
    """Half uniform, half in 16 clusters of sigma 0.02 (the benchmark's distribution)."""
    rng = np.random.default_rng(seed)
    half = n // 2
    centres = rng.uniform(-0.8, 0.8, (2, 16))
    per = (n - half) // 16
    clusters = (centres[:, :, None] + rng.normal(0.0, 0.02, (2, 16, per))).reshape(2, -1)
    rest = n - half - 16 * per
    pos = np.concatenate((rng.uniform(-1, 1, (2, half + rest)), clusters), axis=1).clip(-1, 1)
    return np.ascontiguousarray(pos)


CASES = [
    # name, positions, leaf capacity, theta
    ("cloud_leaf8", cloud(3000, 0), 8, 0.5),
    ("cloud_leaf1", cloud(1500, 1), 1, 0.5),
    ("cloud_leaf16_theta07", cloud(3000, 2), 16, 0.7),
    ("clustered_leaf8_theta03", clustered(5000, 3), 8, 0.3),
]


def _sorted_pairs(pairs):
    """A pair list in lexicographic order, int32 (the lists are sets: order is free)."""
    return pairs[np.lexsort((pairs[:, 1], pairs[:, 0]))].astype(np.int32)


def _compact(a):
    """Integer arrays stored as int32 (every index here is far below 2^31)."""
    return a.astype(np.int32) if a.dtype == np.int64 else a


def main():
    out = {}
    for name, pos, leaf, theta in CASES:
        root = np.array([0.0, 0.0, 1.0])
        t = build(pos, root, leaf)
        m2l, p2p = traverse(t, theta)
        out[f"{name}/pos"] = pos
        out[f"{name}/params"] = np.array([leaf, theta, *root])
        for key, value in t.items():
            out[f"{name}/{key}"] = _compact(value)
        out[f"{name}/m2l"] = _sorted_pairs(m2l)
        out[f"{name}/p2p"] = _sorted_pairs(p2p)
        print(f"{name}: {len(t['node_half'])} nodes, {len(m2l)} M2L, {len(p2p)} P2P pairs")
    # fmm_at's plan: a target grid with its own (non-dyadic) root against a source tree
    rng = np.random.default_rng(4)
    g = np.linspace(-0.93, 0.87, 45)
    gx, gy = np.meshgrid(g, g)
    targets = np.concatenate((np.stack((gx.ravel(), gy.ravel())), rng.uniform(-1, 1, (2, 300))),
                             axis=1)
    sources = cloud(3000, 5)
    for leaf, theta in ((8, 0.5), (4, 0.6)):
        name = f"dual_leaf{leaf}"
        s = build(sources, np.array([0.0, 0.0, 1.0]), 8)
        troot = target_root(targets)
        t = build(targets, troot, leaf)
        m2l, p2p = dual(t, s, theta)
        out[f"{name}/targets"] = targets
        out[f"{name}/sources"] = sources
        out[f"{name}/params"] = np.array([leaf, theta, *troot])
        for key, value in t.items():
            out[f"{name}/t_{key}"] = _compact(value)
        for key, value in s.items():
            out[f"{name}/s_{key}"] = _compact(value)
        out[f"{name}/m2l"] = _sorted_pairs(m2l)
        out[f"{name}/p2p"] = _sorted_pairs(p2p)
        print(f"{name}: {len(t['node_half'])} target nodes, {len(m2l)} M2L, {len(p2p)} P2P")
    np.savez_compressed(os.path.join(HERE, "traversal.npz"), **out)


if __name__ == "__main__":
    main()
