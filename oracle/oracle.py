"""CPU oracle (test infrastructure only): the reference's algorithm restated in C.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, as the checker or as the timed CPU baseline. The
product path (``tree-d-3d_amd/``) never imports it.

The kernels live in ``td3_oracle.c`` (built by ``oracle/Makefile`` into
``libtd3_oracle.so``); this module holds the orchestration of the reference's
``flat_tree.py`` that calls them:

* `direct_field`            flat_tree.py:125-160
* `OracleTree.__init__`     flat_tree.py:746-809 (checks), `_build_host` :483-530
* `OracleTree._moments`     flat_tree.py:663-744 (the CPU upward pass)
* `OracleTree.evaluate`     flat_tree.py:988-1061 with `_evaluate_groups` :914-963 and
                            `_target_order` :891-912
* `OracleTree.traverse`     flat_tree.py:1318-1358
* `OracleTree.fmm`          flat_tree.py:1461-1634 (log kernel, one process)

Pinning: ``tests/test_oracle_*.py`` checks every function here against closed forms,
brute force on tiny inputs and the invariants that fix the tree uniquely.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libtd3_oracle.so")
_LIB = None

MAX_ORDER = 64
GROUP_SIZE = 32
_KERNELS = ("coulomb", "log")

_P = ctypes.c_void_p
_I = ctypes.c_int64
_D = ctypes.c_double
_C = ctypes.c_int


def build() -> str:
    """Compile ``libtd3_oracle.so`` with ``make`` (gcc -O3 -fopenmp)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _LIB = ctypes.CDLL(_LIB_PATH)
        sig = {
            "oracle_direct": [_C, _P, _I, _I, _P, _P, _I, _I, _P, _D, _P, _P, _I, _P],
            "oracle_quadtree_build": [_P, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P],
            "oracle_quadtree_moments_upward": [_P, _I, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P],
            "oracle_binomial_table": [_I, _P],
            "oracle_m2l_coefficients": [_I, _C, _P],
            "oracle_log_moments_upward": [_P, _I, _P, _P, _P, _I, _P, _P, _P, _P, _P, _I, _I, _P, _P, _P],
            "oracle_evaluate_groups": [_P, _I, _P, _P, _I, _C, _C, _P, _I, _P, _P, _P, _P, _P, _P,
                                       _P, _P, _P, _P, _P, _P, _P, _I, _D, _D, _I, _P, _P, _I],
            "oracle_fmm_traverse": [_P, _P, _P, _P, _P, _D, _P, _I, _P, _I, _P, _I, _P, _I, _P],
            "oracle_pair_csr": [_P, _I, _C, _I, _P, _P],
            "oracle_fmm_m2l_host": [_P, _P, _I, _P, _P, _P, _P, _P, _I, _P, _P],
            "oracle_fmm_downward": [_I, _I, _P, _P, _P, _I, _I, _P, _P],
            "oracle_fmm_evaluate_host": [_P, _I, _P, _I, _P, _P, _P, _P, _P, _P, _I, _P, _I, _D, _C,
                                         _P, _P, _I, _P],
            "oracle_dual_traverse": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _D, _P, _I, _P, _I, _P,
                                     _I, _P],
            "oracle_log_targets_host": [_P, _I, _P, _I, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _I,
                                        _P, _P, _P, _P, _D, _C, _P, _P, _I],
            "oracle_cart_moments_upward": [_P, _I, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P, _P,
                                           _I, _I, _P],
            "oracle_cart_m2l_host": [_P, _P, _I, _P, _P, _P, _I, _D, _P, _P, _P, _P, _P, _I, _P],
            "oracle_cart_downward": [_P, _I, _P, _P, _P, _P, _P, _P, _I, _I, _P],
            "oracle_cart_evaluate_host": [_P, _I, _P, _I, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P,
                                          _P, _P, _I, _D, _P, _P, _I, _P],
            "oracle_cart_targets_host": [_P, _I, _P, _I, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P, _P,
                                         _P, _P, _P, _I, _P, _P, _P, _P, _D, _P, _P, _I],
        }
        for name, args in sig.items():
            fn = getattr(_LIB, name)
            fn.argtypes = args
            fn.restype = None
    return _LIB


def _ptr(a: np.ndarray) -> int:
    assert a.flags.c_contiguous, "oracle arrays must be C-contiguous"
    return a.ctypes.data


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))


def _coincident() -> ValueError:
    return ValueError("Coincident charges require positive softening")


def _check(status) -> None:
    """_utils.py:26-35."""
    code = int(status[0])
    if code > 0:
        raise _coincident()
    if code < 0:
        raise RuntimeError("the quadtree is too deep for the traversal stack")


def _points(array, name: str) -> np.ndarray:
    """_utils.py:51-63: a contiguous (2, N) float64 array."""
    points = _f64(array)
    if points.ndim != 2 or points.shape[0] != 2:
        raise ValueError(f"{name} must have shape (2, N) (component-major), got {points.shape}")
    return points.copy(order="C")


def binomial_table(size: int) -> np.ndarray:
    out = np.zeros((size, size))
    lib().oracle_binomial_table(size, _ptr(out))
    return out


def m2l_coefficients(order: int, symmetric: bool) -> np.ndarray:
    out = np.zeros((order + 1, order + 1))
    lib().oracle_m2l_coefficients(order, 1 if symmetric else 0, _ptr(out))
    return out


def direct_field(targets, pos, charge, *, softening=0.0, exclude=None, kernel="coulomb"):
    """flat_tree.py:125-160 on the CPU oracle."""
    if kernel not in _KERNELS:
        raise ValueError(f"kernel must be one of {_KERNELS}")
    targets = _points(targets, "targets")
    pos = _points(pos, "pos")
    charge = _f64(charge).reshape(-1)
    m, n = targets.shape[1], pos.shape[1]
    exclude = np.full(m, -1, dtype=np.int64) if exclude is None else _i64(exclude).reshape(-1)
    if not m or not n:
        return np.zeros(m), np.zeros((2, m))
    phi, field = np.zeros(m), np.zeros((2, m))
    status = np.zeros(1, dtype=np.int64)
    lib().oracle_direct(1 if kernel == "log" else 0, _ptr(targets), m, m, _ptr(exclude),
                        _ptr(pos), n, n, _ptr(charge), float(softening), _ptr(phi), _ptr(field),
                        m, _ptr(status))
    _check(status)
    return phi, field


class OracleTree:
    """The reference's FlatQuadtree on the CPU oracle (one process, FP64)."""

    def __init__(self, positions, charges, root, *, leaf_capacity=8, max_depth=48,
                 kernel="coulomb", order=None, symmetric=False):
        if kernel not in _KERNELS:
            raise ValueError(f"kernel must be one of {_KERNELS}")
        if order is None:
            order = 12 if kernel == "log" else 4
        if int(order) != order or order < 1 or order > MAX_ORDER:
            raise ValueError("bad order")
        self.kernel, self.order, self.symmetric = kernel, int(order), bool(symmetric)
        self.leaf_capacity, self.max_depth = int(leaf_capacity), int(max_depth)
        xmid, ymid, half = (float(v) for v in root)
        self.pos = _points(positions, "positions")
        self.charge = _f64(charges).reshape(-1).copy()
        n = self.pos.shape[1]
        if len(self.charge) != n:
            raise ValueError("positions and charges must have the same length")
        if not np.all(np.isfinite(self.pos)):
            raise ValueError("Particle position must be finite")
        if np.any(self.pos[0] < xmid - half) or np.any(self.pos[0] > xmid + half) or \
           np.any(self.pos[1] < ymid - half) or np.any(self.pos[1] > ymid + half):
            raise ValueError("Particle lies outside tree bounds")
        self._traversals = {}
        if not n:
            self.perm = np.empty(0, dtype=np.int64)
            self.node_count = 1
            self.node_center = np.array([[xmid, ymid]])
            self.node_half = np.array([half])
            self.node_start = np.zeros(1, dtype=np.int64)
            self.node_end = np.zeros(1, dtype=np.int64)
            self.node_child = np.full(1, -1, dtype=np.int64)
            self.node_depth = np.zeros(1, dtype=np.int64)
            self.levels = np.array([0, 1])
            self.node_q = np.zeros(1)
            self.node_qabs = np.zeros(1)
            self.node_p = np.zeros((1, 2))
            self.node_s = np.zeros((1, 3))
            self.moment_re = np.zeros((1, self.order + 1))
            self.moment_im = np.zeros((1, self.order + 1))
            return
        self._build(xmid, ymid, half, 4 * (2 * n // self.leaf_capacity + 16))
        self._moments()

    # flat_tree.py:483-530
    def _build(self, xmid, ymid, half, capacity):
        n = self.pos.shape[1]
        sizes = np.zeros(3, dtype=np.int64)
        self.build_calls = 0
        while True:
            self.build_calls += 1
            sorted_pos = np.zeros((2, n + 1))
            sorted_pos[:, :n] = self.pos
            center = np.empty((capacity, 2))
            half_arr = np.empty(capacity)
            start = np.empty(capacity, dtype=np.int64)
            end = np.empty(capacity, dtype=np.int64)
            child = np.empty(capacity, dtype=np.int64)
            depth = np.empty(capacity, dtype=np.int64)
            center[0] = (xmid, ymid)
            half_arr[0] = half
            perm = np.empty(n, dtype=np.int64)
            levels = np.empty(capacity + 1, dtype=np.int64)
            lib().oracle_quadtree_build(_ptr(sorted_pos), n + 1, n, self.leaf_capacity,
                                        self.max_depth, _ptr(center), _ptr(half_arr), _ptr(start),
                                        _ptr(end), _ptr(child), _ptr(depth), _ptr(perm),
                                        _ptr(levels), capacity, _ptr(sizes))
            if not sizes[2]:
                break
            capacity *= 2
        count = int(sizes[0])
        self.node_count = count
        self.node_center = center[:count].copy()
        self.node_half = half_arr[:count].copy()
        self.node_start = start[:count].copy()
        self.node_end = end[:count].copy()
        self.node_child = child[:count].copy()
        self.node_depth = depth[:count].copy()
        self.perm = perm
        self.levels = levels[: sizes[1] + 1].copy()
        self._sorted_pos = sorted_pos

    # flat_tree.py:663-744 (CPU, one process: the upward pass)
    def _moments(self):
        c = self.node_count
        lev = _i64(self.levels)
        nl = len(lev) - 1
        n = self.pos.shape[1]
        self.node_q = np.empty(c)
        self.node_qabs = np.empty(c)
        self.node_p = np.empty((c, 2))
        self.node_s = np.empty((c, 3))
        if self.kernel == "log":
            p1 = self.order + 1
            self.moment_re = np.empty((c, p1))
            self.moment_im = np.empty((c, p1))
            bsz = 2 * self.order + 1
            binom = binomial_table(bsz)
            lib().oracle_log_moments_upward(
                _ptr(self.pos), n, _ptr(self.charge), _ptr(self.perm), _ptr(lev), nl,
                _ptr(self.node_center), _ptr(self.node_start), _ptr(self.node_end),
                _ptr(self.node_child), _ptr(binom), bsz, self.order, _ptr(self.node_qabs),
                _ptr(self.moment_re), _ptr(self.moment_im))
            self.node_q[:] = self.moment_re[:, 0]
            self.node_p[:, 0] = self.moment_re[:, 1]
            self.node_p[:, 1] = self.moment_im[:, 1]
            return
        lib().oracle_quadtree_moments_upward(
            _ptr(self.pos), n, _ptr(self.charge), _ptr(self.perm), _ptr(lev), nl,
            _ptr(self.node_center), _ptr(self.node_start), _ptr(self.node_end),
            _ptr(self.node_child), _ptr(self.node_q), _ptr(self.node_qabs), _ptr(self.node_p),
            _ptr(self.node_s))
        self.moment_re = np.zeros((1, 1))
        self.moment_im = np.zeros((1, 1))

    def set_charges(self, charges):
        """flat_tree.py:853-868."""
        charges = _f64(charges).reshape(-1)
        if charges.shape[0] != self.pos.shape[1]:
            raise ValueError("charges must have one entry per particle")
        self.charge[...] = charges
        if self.pos.shape[1]:
            self._moments()

    def _sorted_sources(self):
        """flat_tree.py:879-889."""
        n = self.pos.shape[1]
        spos = np.zeros((2, n + 1))
        spos[:, :n] = self.pos[:, self.perm]
        return spos, np.ascontiguousarray(self.charge[self.perm])

    def _target_order(self, targets):
        """flat_tree.py:891-912: 16-bit-per-axis Morton order (locality only)."""
        if targets.shape == self.pos.shape and np.array_equal(targets, self.pos):
            return self.perm
        lo = targets.min(axis=1, keepdims=True)
        span = float(np.max(targets.max(axis=1, keepdims=True) - lo))
        cells = ((targets - lo) * ((65535.0 / span) if span > 0 else 0.0)).astype(np.uint64)
        key = np.zeros(targets.shape[1], dtype=np.uint64)
        for axis, shift in ((0, 1), (1, 0)):
            v = cells[axis]
            v = (v | (v << np.uint64(8))) & np.uint64(0x00FF00FF)
            v = (v | (v << np.uint64(4))) & np.uint64(0x0F0F0F0F)
            v = (v | (v << np.uint64(2))) & np.uint64(0x33333333)
            v = (v | (v << np.uint64(1))) & np.uint64(0x55555555)
            key |= v << np.uint64(shift)
        return np.argsort(key, kind="stable")

    def evaluate(self, targets, *, exclude=None, theta=0.5, softening=0.0, potential=True):
        """flat_tree.py:988-1061 with `_evaluate_groups` :914-963."""
        targets = _points(targets, "targets")
        m = targets.shape[1]
        exclude = np.full(m, -1, dtype=np.int64) if exclude is None else _i64(exclude).reshape(-1)
        phi, field = np.zeros(m), np.zeros((2, m))
        n = self.pos.shape[1]
        if not (m and n):
            return phi, field
        order = self._target_order(targets)
        spos, scharge = self._sorted_sources()
        log_kernel = self.kernel == "log"
        st = np.empty((2, m + 1))
        st[:, :m] = targets[:, order]
        phi_s = np.empty(m)
        field_s = np.empty((2, m + 1))
        groups = _i64(np.append(np.arange(0, m, GROUP_SIZE), m))
        excl = _i64(exclude[order])
        stack_size = 3 * int(self.node_depth.max()) + 8
        mre = self.moment_re if log_kernel else np.zeros((1, 1))
        mim = self.moment_im if log_kernel else np.zeros((1, 1))
        lib().oracle_evaluate_groups(
            _ptr(st), m + 1, _ptr(excl), _ptr(groups), len(groups) - 1, 1 if log_kernel else 0,
            1 if potential else 0, _ptr(spos), n + 1, _ptr(scharge), _ptr(self.perm),
            _ptr(self.node_center), _ptr(self.node_half), _ptr(self.node_start),
            _ptr(self.node_end), _ptr(self.node_child), _ptr(self.node_q), _ptr(self.node_qabs),
            _ptr(self.node_p), _ptr(self.node_s), _ptr(_f64(mre)), _ptr(_f64(mim)),
            self.order if log_kernel else 0, float(theta), float(softening), stack_size,
            _ptr(phi_s), _ptr(field_s), m + 1)
        phi[order] = phi_s
        field[:, order] = field_s[:, :m]
        bad = np.flatnonzero(phi == 1.0e300)
        if bad.size:
            raise _coincident()
        if not potential:
            phi[...] = 0.0
        return phi, field

    def traverse(self, theta):
        """flat_tree.py:1318-1358: generations of fmm_traverse from (0, 0)."""
        if theta in self._traversals:
            return self._traversals[theta]
        c = self.node_count
        capacity = np.array([16, 16, 10]) * c + 64
        self.traverse_restarts = 0
        while True:
            m2l = np.empty((capacity[0], 2), dtype=np.int64)
            p2p = np.empty((capacity[1], 2), dtype=np.int64)
            frontier = np.zeros((capacity[2], 2), dtype=np.int64)
            nxt = np.empty((capacity[2], 2), dtype=np.int64)
            counts = np.zeros(4, dtype=np.int64)
            n_frontier = 1
            while n_frontier:
                counts[2] = 0
                lib().oracle_fmm_traverse(
                    _ptr(self.node_center), _ptr(self.node_half), _ptr(self.node_start),
                    _ptr(self.node_end), _ptr(self.node_child), float(theta), _ptr(frontier),
                    n_frontier, _ptr(nxt), capacity[2], _ptr(m2l), capacity[0], _ptr(p2p),
                    capacity[1], _ptr(counts))
                if counts[3]:
                    break
                n_frontier = int(counts[2])
                frontier, nxt = nxt, frontier
            if not counts[3]:
                out = (m2l[: counts[0]].copy(), p2p[: counts[1]].copy())
                self._traversals[theta] = out
                return out
            self.traverse_restarts += 1
            capacity = np.maximum(2 * capacity, counts[:3])

    def fmm(self, *, theta=0.5, softening=0.0, potential=True):
        """flat_tree.py:1461-1634 for the log kernel, one process (CPU branch)."""
        if not 0 < theta < 1:
            raise ValueError("FMM theta must lie in (0, 1)")
        if self.kernel != "log":
            raise NotImplementedError("the oracle's FMM covers the log kernel")
        n = self.pos.shape[1]
        phi, field = np.zeros(n), np.zeros((2, n))
        if not n:
            return phi, field
        m2l, p2p = self.traverse(theta)
        c, order = self.node_count, self.order
        p1 = order + 1
        local_re = np.zeros((c, p1))
        local_im = np.zeros((c, p1))
        if len(m2l):
            rows = np.zeros(c + 1, dtype=np.int64)
            partners = np.empty(max(2 * len(m2l), 1), dtype=np.int64)
            lib().oracle_pair_csr(_ptr(_i64(m2l)), len(m2l), 1, c, _ptr(rows), _ptr(partners))
            coeff = m2l_coefficients(order, self.symmetric)
            lib().oracle_fmm_m2l_host(_ptr(rows), _ptr(partners), c, _ptr(self.node_center),
                                      _ptr(self.node_center), _ptr(self.moment_re),
                                      _ptr(self.moment_im), _ptr(coeff), order, _ptr(local_re),
                                      _ptr(local_im))
        bsz = 2 * order + 1
        binom = binomial_table(bsz)
        for lo, hi in zip(self.levels[:-2], self.levels[1:-1]):
            lib().oracle_fmm_downward(int(lo), int(hi), _ptr(self.node_center),
                                      _ptr(self.node_child), _ptr(binom), bsz, order,
                                      _ptr(local_re), _ptr(local_im))
        self.local_re, self.local_im = local_re, local_im
        spos, scharge = self._sorted_sources()
        phi_s = np.zeros(n)
        field_s = np.zeros((2, n + 1))
        status = np.zeros(1, dtype=np.int64)
        pairs = _i64(p2p) if len(p2p) else np.zeros((1, 2), dtype=np.int64)
        lib().oracle_fmm_evaluate_host(
            _ptr(spos), n + 1, _ptr(scharge), c, _ptr(self.node_center), _ptr(self.node_start),
            _ptr(self.node_end), _ptr(self.node_child), _ptr(local_re), _ptr(local_im), order,
            _ptr(pairs), len(p2p), float(softening), 1 if potential else 0, _ptr(phi_s),
            _ptr(field_s), n + 1, _ptr(status))
        if int(status[0]):
            status[0] = self.perm[int(status[0]) - 1] + 1
        _check(status)
        phi[self.perm] = phi_s
        field[:, self.perm] = field_s[:, :n]
        self.fmm_pairs = (len(m2l), len(p2p))
        if not potential:
            phi[:] = 0.0
        return phi, field


def _fmm_at_log(self, targets, *, exclude=None, theta=0.5, softening=0.0, potential=True,
                leaf_capacity=None):
    """flat_tree.py:1152-1316 for the log kernel: target tree, dual traversal, M2L into the
    target cells, L2L down the target tree, log_targets_host."""
    if not 0 < theta < 1:
        raise ValueError("FMM theta must lie in (0, 1)")
    targets = _points(targets, "targets")
    if not np.all(np.isfinite(targets)):
        raise ValueError("targets must be finite")
    m, n = targets.shape[1], self.pos.shape[1]
    phi, field = np.zeros(m), np.zeros((2, m))
    if not m or not n:
        return phi, field
    exclude = np.full(m, -1, dtype=np.int64) if exclude is None else _i64(exclude).reshape(-1)
    leaf = self.leaf_capacity if leaf_capacity is None else int(leaf_capacity)
    lo, hi = targets.min(axis=1), targets.max(axis=1)
    centre = 0.5 * (lo + hi)
    half = 0.5 * float(np.max(hi - lo)) * (1 + 1e-12) + 1e-300
    tree = OracleTree(targets, np.zeros(m), (float(centre[0]), float(centre[1]), half),
                      leaf_capacity=leaf, max_depth=self.max_depth, kernel="coulomb")
    depth = int(tree.node_depth.max()) + int(self.node_depth.max())
    cap_stack = 3 * depth + 8
    capacity = np.array([16, 8]) * (tree.node_count + self.node_count) + 64
    counts = np.zeros(3, dtype=np.int64)
    while True:
        stack = np.empty((cap_stack, 2), dtype=np.int64)
        m2l = np.empty((capacity[0], 2), dtype=np.int64)
        p2p = np.empty((capacity[1], 2), dtype=np.int64)
        lib().oracle_dual_traverse(
            _ptr(tree.node_center), _ptr(tree.node_half), _ptr(tree.node_start),
            _ptr(tree.node_end), _ptr(tree.node_child), _ptr(self.node_center),
            _ptr(self.node_half), _ptr(self.node_start), _ptr(self.node_end),
            _ptr(self.node_child), float(theta), _ptr(stack), cap_stack, _ptr(m2l), capacity[0],
            _ptr(p2p), capacity[1], _ptr(counts))
        if not counts[2]:
            break
        capacity = np.maximum(2 * capacity, counts[:2] + 64)
    m2l, p2p = m2l[: counts[0]].copy(), p2p[: counts[1]].copy()
    self.dual_lists = (m2l, p2p)
    ct, order = tree.node_count, self.order
    if self.kernel != "log":
        return _fmm_at_coulomb(self, tree, m2l, p2p, targets, exclude, softening, potential)
    rows = []
    for pairs in (m2l, p2p):
        start = np.zeros(ct + 1, dtype=np.int64)
        partners = np.empty(max(len(pairs), 1), dtype=np.int64)
        lib().oracle_pair_csr(_ptr(_i64(pairs) if len(pairs) else np.zeros((1, 2), np.int64)),
                              len(pairs), 0, ct, _ptr(start), _ptr(partners))
        rows.append((start, partners))
    p1 = order + 1
    local_re = np.zeros((ct, p1))
    local_im = np.zeros((ct, p1))
    if len(m2l):
        coeff = m2l_coefficients(order, self.symmetric)
        lib().oracle_fmm_m2l_host(_ptr(rows[0][0]), _ptr(rows[0][1]), ct, _ptr(tree.node_center),
                                  _ptr(self.node_center), _ptr(self.moment_re),
                                  _ptr(self.moment_im), _ptr(coeff), order, _ptr(local_re),
                                  _ptr(local_im))
    bsz = 2 * order + 1
    binom = binomial_table(bsz)
    for a, b in zip(tree.levels[:-2], tree.levels[1:-1]):
        lib().oracle_fmm_downward(int(a), int(b), _ptr(tree.node_center), _ptr(tree.node_child),
                                  _ptr(binom), bsz, order, _ptr(local_re), _ptr(local_im))
    tpos, _ = tree._sorted_sources()
    texclude = _i64(exclude[tree.perm])
    spos, scharge = self._sorted_sources()
    phi_s = np.empty(m)
    field_s = np.empty((2, m + 1))
    lib().oracle_log_targets_host(
        _ptr(tpos), m + 1, _ptr(texclude), ct, _ptr(tree.node_center), _ptr(tree.node_start),
        _ptr(tree.node_end), _ptr(tree.node_child), _ptr(local_re), _ptr(local_im), order,
        _ptr(rows[1][0]), _ptr(rows[1][1]), _ptr(spos), n + 1, _ptr(scharge), _ptr(self.perm),
        _ptr(self.node_start), _ptr(self.node_end), float(softening), 1 if potential else 0,
        _ptr(phi_s), _ptr(field_s), m + 1)
    phi[tree.perm] = phi_s
    field[:, tree.perm] = field_s[:, :m]
    if np.any(phi == 1.0e300):
        raise _coincident()
    if not potential:
        phi[:] = 0.0
    return phi, field


OracleTree.fmm_at = _fmm_at_log


# ------------------------------------------------------------------ the Coulomb FMM
_CART: dict = {}


def cart_size(order: int) -> int:
    """_cartesian.py:29-31."""
    return (order + 1) * (order + 2) // 2


def cart_index(a: int, b: int) -> int:
    """_cartesian.py:34-37: coefficient (a, b) at n(n+1)/2 + b, n = a + b."""
    n = a + b
    return n * (n + 1) // 2 + b


def cart_tables(order: int) -> dict:
    """_cartesian.py:40-107: the index tables of the Cartesian expansions."""
    from math import comb

    if order in _CART:
        return _CART[order]
    nc = cart_size(order)
    ex = np.empty(nc, dtype=np.int64)
    ey = np.empty(nc, dtype=np.int64)
    for n in range(order + 1):
        for b in range(n + 1):
            ex[cart_index(n - b, b)] = n - b
            ey[cart_index(n - b, b)] = b

    def at(a, b):
        return cart_index(a, b) if a >= 0 and b >= 0 else -1

    mono_parent = np.zeros(nc, dtype=np.int64)
    mono_dir = np.zeros(nc, dtype=np.int64)
    rec = np.full((nc, 4), -1, dtype=np.int64)
    grad = np.full((nc, 2), -1, dtype=np.int64)
    for i in range(nc):
        a, b = int(ex[i]), int(ey[i])
        if a > 0:
            mono_parent[i], mono_dir[i] = at(a - 1, b), 0
        elif b > 0:
            mono_parent[i], mono_dir[i] = at(a, b - 1), 1
        rec[i] = (at(a - 1, b), at(a, b - 1), at(a - 2, b), at(a, b - 2))
        grad[i] = (at(a - 1, b), at(a, b - 1))
    m2l, m2l_coef, shift, shift_coef = [], [], [], []
    for ib in range(nc):
        bx, by = int(ex[ib]), int(ey[ib])
        for ia in range(cart_size(order - bx - by)):
            ax, ay = int(ex[ia]), int(ey[ia])
            m2l.append((ib, ia, cart_index(ax + bx, ay + by)))
            m2l_coef.append((-1) ** (ax + ay) * comb(ax + bx, ax) * comb(ay + by, ay))
    for ig in range(nc):
        gx, gy = int(ex[ig]), int(ey[ig])
        for bx in range(gx + 1):
            for by in range(gy + 1):
                shift.append((ig, cart_index(bx, by), cart_index(gx - bx, gy - by)))
                shift_coef.append(comb(gx, bx) * comb(gy, by))
    out = dict(ex=ex, ey=ey, mono_parent=mono_parent, mono_dir=mono_dir, rec=rec, grad=grad,
               m2l=np.array(m2l, dtype=np.int64), m2l_coef=np.array(m2l_coef, dtype=float),
               shift=np.array(shift, dtype=np.int64), shift_coef=np.array(shift_coef, dtype=float))
    _CART[order] = out
    return out


def _cart_moments(self):
    """flat_tree.py:1063-1084 (_cartesian_moments)."""
    tb = cart_tables(self.order)
    nc = cart_size(self.order)
    moments = np.empty((self.node_count, nc))
    lev = _i64(self.levels)
    lib().oracle_cart_moments_upward(
        _ptr(self.pos), self.pos.shape[1], _ptr(self.charge), _ptr(self.perm), _ptr(lev),
        len(lev) - 1, _ptr(self.node_center), _ptr(self.node_start), _ptr(self.node_end),
        _ptr(self.node_child), _ptr(tb["mono_parent"]), _ptr(tb["mono_dir"]), _ptr(tb["shift"]),
        _ptr(tb["shift_coef"]), len(tb["shift"]), nc, _ptr(moments))
    return moments


def _cart_local(self, dst_tree, rows, partners, softening):
    tb = cart_tables(self.order)
    nc = cart_size(self.order)
    local = np.zeros((dst_tree.node_count, nc))
    moments = _cart_moments(self)
    lib().oracle_cart_m2l_host(
        _ptr(rows), _ptr(partners), dst_tree.node_count, _ptr(dst_tree.node_center),
        _ptr(self.node_center), _ptr(moments), nc, float(softening), _ptr(tb["ex"]),
        _ptr(tb["ey"]), _ptr(tb["rec"]), _ptr(tb["m2l"]), _ptr(tb["m2l_coef"]), len(tb["m2l"]),
        _ptr(local))
    lev = _i64(dst_tree.levels)
    lib().oracle_cart_downward(
        _ptr(lev), len(lev) - 1, _ptr(dst_tree.node_center), _ptr(dst_tree.node_child),
        _ptr(tb["mono_parent"]), _ptr(tb["mono_dir"]), _ptr(tb["shift"]), _ptr(tb["shift_coef"]),
        len(tb["shift"]), nc, _ptr(local))
    return local


def _coulomb_fmm(self, *, theta=0.5, softening=0.0, potential=True):
    """flat_tree.py:1360-1428 (_coulomb_fmm, CPU)."""
    if not 0 < theta < 1:
        raise ValueError("FMM theta must lie in (0, 1)")
    n = self.pos.shape[1]
    phi, field = np.zeros(n), np.zeros((2, n))
    if not n:
        return phi, field
    m2l, p2p = self.traverse(theta)
    c = self.node_count
    rows = np.zeros(c + 1, dtype=np.int64)
    partners = np.empty(max(2 * len(m2l), 1), dtype=np.int64)
    if len(m2l):
        lib().oracle_pair_csr(_ptr(_i64(m2l)), len(m2l), 1, c, _ptr(rows), _ptr(partners))
    local = _cart_local(self, self, rows, partners, softening)
    self.cart_local = local
    tb = cart_tables(self.order)
    nc = cart_size(self.order)
    spos, scharge = self._sorted_sources()
    phi_s = np.zeros(n)
    field_s = np.zeros((2, n + 1))
    status = np.zeros(1, dtype=np.int64)
    pairs = _i64(p2p) if len(p2p) else np.zeros((1, 2), dtype=np.int64)
    lib().oracle_cart_evaluate_host(
        _ptr(spos), n + 1, _ptr(scharge), c, _ptr(self.node_center), _ptr(self.node_start),
        _ptr(self.node_end), _ptr(self.node_child), _ptr(local), nc, _ptr(tb["mono_parent"]),
        _ptr(tb["mono_dir"]), _ptr(tb["ex"]), _ptr(tb["ey"]), _ptr(tb["grad"]), _ptr(pairs),
        len(p2p), float(softening), _ptr(phi_s), _ptr(field_s), n + 1, _ptr(status))
    if int(status[0]):
        status[0] = self.perm[int(status[0]) - 1] + 1
    _check(status)
    phi[self.perm] = phi_s
    field[:, self.perm] = field_s[:, :n]
    self.fmm_pairs = (len(m2l), len(p2p))
    return phi, field


_log_fmm = OracleTree.fmm


def _fmm(self, *, theta=0.5, softening=0.0, potential=True):
    if self.kernel == "log":
        return _log_fmm(self, theta=theta, softening=softening, potential=potential)
    return _coulomb_fmm(self, theta=theta, softening=softening, potential=potential)


OracleTree.fmm = _fmm


def _fmm_at_coulomb(self, tree, m2l, p2p, targets, exclude, softening, potential):
    """fmm_at's Coulomb branch (flat_tree.py:1247-1295): cart_m2l_host, cart_downward,
    cart_targets_host."""
    m, n = targets.shape[1], self.pos.shape[1]
    ct = tree.node_count
    rows = []
    for pairs in (m2l, p2p):
        start = np.zeros(ct + 1, dtype=np.int64)
        partners = np.empty(max(len(pairs), 1), dtype=np.int64)
        lib().oracle_pair_csr(_ptr(_i64(pairs) if len(pairs) else np.zeros((1, 2), np.int64)),
                              len(pairs), 0, ct, _ptr(start), _ptr(partners))
        rows.append((start, partners))
    local = _cart_local(self, tree, rows[0][0], rows[0][1], softening)
    tb = cart_tables(self.order)
    nc = cart_size(self.order)
    tpos, _ = tree._sorted_sources()
    texclude = _i64(exclude[tree.perm])
    spos, scharge = self._sorted_sources()
    phi_s = np.empty(m)
    field_s = np.empty((2, m + 1))
    lib().oracle_cart_targets_host(
        _ptr(tpos), m + 1, _ptr(texclude), ct, _ptr(tree.node_center), _ptr(tree.node_start),
        _ptr(tree.node_end), _ptr(tree.node_child), _ptr(local), nc, _ptr(tb["mono_parent"]),
        _ptr(tb["mono_dir"]), _ptr(tb["ex"]), _ptr(tb["ey"]), _ptr(tb["grad"]),
        _ptr(rows[1][0]), _ptr(rows[1][1]), _ptr(spos), n + 1, _ptr(scharge), _ptr(self.perm),
        _ptr(self.node_start), _ptr(self.node_end), float(softening), _ptr(phi_s),
        _ptr(field_s), m + 1)
    phi, field = np.zeros(m), np.zeros((2, m))
    phi[tree.perm] = phi_s
    field[:, tree.perm] = field_s[:, :m]
    if np.any(phi == 1.0e300):
        raise _coincident()
    if not potential:
        phi[:] = 0.0
    return phi, field


def push(positions, velocities, forces, electric, masses, charges, fixed, dt, b_z, first_half):
    """The kick-drift (``first_half``) or closing kick of a step, in place on NumPy arrays
    (simulation.py:812-835): half-kick of mobile particles by ``electric`` (qE), then the
    exact cyclotron drift and rotation by alpha = (q/m) B_z dt; the closing half sets
    ``forces = electric + q B_z (v_y, -v_x)`` for every particle."""
    mobile = ~np.asarray(fixed, dtype=bool)
    inv_mass = 1 / masses[mobile]
    velocities[:, mobile] += 0.5 * dt * electric[:, mobile] * inv_mass
    if first_half:
        angle = charges[mobile] / masses[mobile] * b_z * dt
        c, s = np.cos(angle), np.sin(angle)
        a = np.sinc(angle / np.pi)
        b = 0.5 * angle * np.sinc(angle / (2 * np.pi)) ** 2
        vx, vy = velocities[0, mobile], velocities[1, mobile]
        positions[:, mobile] += dt * np.stack((a * vx + b * vy, a * vy - b * vx))
        velocities[:, mobile] = np.stack((c * vx + s * vy, c * vy - s * vx))
    else:
        forces[:] = electric + charges * b_z * np.stack((velocities[1], -velocities[0]))


# ------------------------------------------------------------------ the grounded strip
def soften_pairs(pos, charge, pairs, eps2, with_potential, phi, field):
    """kernels/soften_pairs/soften_pairs_pyccel.py:14-60, in place: for each listed pair the
    softened minus bare line-charge kernel added to both partners; raises on a zero
    separation."""
    for i, j in pairs:
        dx = pos[0, i] - pos[0, j]
        dy = pos[1, i] - pos[1, j]
        r2 = dx * dx + dy * dy
        if r2 == 0.0:
            raise ValueError("Coincident charges")
        soft = r2 + eps2
        if with_potential:
            dphi = -np.log(soft / r2)
            phi[i] += charge[j] * dphi
            phi[j] += charge[i] * dphi
        scale = 2.0 / soft - 2.0 / r2
        field[0, i] += charge[j] * scale * dx
        field[1, i] += charge[j] * scale * dy
        field[0, j] -= charge[i] * scale * dx
        field[1, j] -= charge[i] * scale * dy


def _strip_map(width, positions):
    """strip.py:136-141: zeta = exp(i pi z / L), charges strictly between the plates."""
    xy = np.asarray(positions, dtype=float).reshape(2, -1)
    if np.any((xy[0] <= 0) | (xy[0] >= width)):
        raise ValueError("Charges must lie strictly between the plates")
    return np.exp(1j * np.pi * (xy[0] + 1j * xy[1]) / width)


def _strip_root(xy):
    """strip.py:153-158."""
    lo, hi = xy.min(axis=1), xy.max(axis=1)
    centre = (lo + hi) / 2
    half = float(np.max(hi - lo) / 2) * (1 + 1e-9) + 1e-300
    return (float(centre[0]), float(centre[1]), half)


def strip_field_at_charges(width, positions, charges, *, order=14, theta=0.5, leaf_capacity=32,
                           k_e=8.9875517923e9, method="fmm", softening=0.0, cutoff=6.0,
                           potential=True):
    """strip.py:185-233 (GroundedStripField.field_at_charges): the charges and their images in
    the mapped plane, one FMM (or direct sum), mapped back with each charge's regular
    self term, then the close pairs softened in the physical plane (k-d tree pairs within
    cutoff * softening, strip.py:168-180)."""
    from scipy.spatial import cKDTree

    zeta = _strip_map(width, positions)
    q0 = np.asarray(charges, dtype=float).ravel()
    n = len(zeta)
    if not n:
        return np.zeros(0), np.zeros((2, 0))
    points = np.concatenate((zeta, np.conj(zeta)))
    xy = np.ascontiguousarray(np.stack((points.real, points.imag)))
    q = np.concatenate((q0, -q0))
    if method == "fmm":
        tree = OracleTree(xy, q, _strip_root(xy), kernel="log", order=order,
                          leaf_capacity=leaf_capacity)
        phi, raw = tree.fmm(theta=theta, potential=potential)
    else:
        phi, raw = direct_field(xy, xy, q, exclude=np.arange(2 * n), kernel="log")
    mapped = raw[0, :n] - 1j * raw[1, :n]
    complex_field = mapped * (1j * np.pi / width) * zeta
    complex_field += 2 * q0 * (1j * np.pi / (2 * width))
    field = np.stack((complex_field.real, -complex_field.imag), axis=0)
    if potential:
        phi_own = phi[:n] - 2 * q0 * np.log(np.pi * np.abs(zeta) / width)
    else:
        phi_own = np.zeros(n)
    if softening > 0:
        pos = np.ascontiguousarray(positions, dtype=float).reshape(2, -1)
        pairs = cKDTree(pos.T).query_pairs(cutoff * softening, output_type="ndarray")
        soften_pairs(pos, q0, pairs, softening ** 2, potential, phi_own, field)
    return k_e * phi_own, k_e * field


def strip_probe(width, positions, charges, points, *, order=14, theta=0.5, leaf_capacity=32,
                k_e=8.9875517923e9, method="fmm", softening=0.0, cutoff=6.0):
    """strip.py:241-323 (GroundedStripField.probe): the walk in the mapped plane at the
    mapped probes, mapped back; sources within cutoff * softening of a probe softened."""
    from scipy.spatial import cKDTree

    zeta = _strip_map(width, positions)
    q0 = np.asarray(charges, dtype=float).ravel()
    points_xy = np.asarray(points, dtype=float).reshape(2, -1)
    probes = _strip_map(width, points_xy)
    probe_xy = np.ascontiguousarray(np.stack((probes.real, probes.imag)))
    src = np.concatenate((zeta, np.conj(zeta)))
    xy = np.ascontiguousarray(np.stack((src.real, src.imag)))
    q = np.concatenate((q0, -q0))
    if method == "fmm":
        root = _strip_root(np.concatenate((xy, probe_xy), axis=1))
        tree = OracleTree(xy, q, root, kernel="log", order=order, leaf_capacity=leaf_capacity)
        phi, raw = tree.evaluate(probe_xy, theta=theta)
    else:
        phi, raw = direct_field(probe_xy, xy, q, kernel="log")
    mapped = raw[0] - 1j * raw[1]
    complex_field = mapped * (1j * np.pi / width) * probes
    field = np.stack((complex_field.real, -complex_field.imag), axis=0)
    if softening > 0:
        sources = np.asarray(positions, dtype=float).reshape(2, -1).T
        targets = points_xy.T
        near = cKDTree(sources).query_ball_point(targets, cutoff * softening,
                                                 return_sorted=False)
        eps2 = softening ** 2
        for t, lst in enumerate(near):
            for s in lst:
                d = targets[t] - sources[s]
                r2 = float(d @ d)
                if r2 == 0:
                    raise ValueError("A probe coincides with a charge")
                phi[t] -= q0[s] * np.log((r2 + eps2) / r2)
                field[:, t] += q0[s] * (2 / (r2 + eps2) - 2 / r2) * d
    return k_e * phi, k_e * field
