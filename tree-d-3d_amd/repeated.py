"""Repeated sums over one fixed geometry: many charge vectors, the same points.

The reference's boundary-integral electrodes apply their single-layer operator once per
GMRES iteration (electrodes.py:1205-1235): the points never move, so it builds one
log-kernel `FlatQuadtree` per geometry, gives it new charges with `set_charges` for every
product (cells and interaction lists kept, moments recomputed) and sums at the same
targets with `fmm_at`, whose target tree and dual lists are kept between products
(flat_tree.py:1124-1149).  A sum at few targets walks the tree instead
(electrodes.py:267-271, `_sum_at`).  The constants are the reference's
(electrodes.py:70-79): theta 0.5, 16 sources per leaf, the walk at theta 0.7 when the
targets are fewer than the sources over 8, order 32.

`RepeatedSums` is that call pattern on the GPU: the first product builds the source tree
and the target plan, every later one runs only the moments, M2L, L2L and the evaluation
(`plan_builds` and `tree_builds` count the builds, so reuse can be checked).
`as_linear_operator` wraps it for SciPy's iterative solvers, as the reference's
`_Operator` does (electrodes.py:275-285).
"""

from __future__ import annotations

import torch

from ._utils import _device, _out, _points
from .flat_tree import FlatQuadtree
from .quad import Quad

#: electrodes.py:70-79
FMM_AT_THETA = 0.5
FMM_LEAF = 16
FMM_AT_RATIO = 8
WALK_THETA = 0.7
FMM_ORDER = 32


class RepeatedSums:
    """Log-kernel potentials (and fields) at fixed ``targets`` of charges at fixed
    ``sources``, for one charge vector after another.

    ``exclude[i]`` is a source left out of target ``i``'s sum (its own point when the
    targets are some of the sources, as the reference's panel midpoints are).  Each call
    ``sums(charges)`` returns the raw potential ``(M,)`` and field ``(2, M)`` (no k_e)."""

    def __init__(self, sources, targets, root: Quad, *, exclude=None, order: int = FMM_ORDER,
                 theta: float = FMM_AT_THETA, leaf_capacity: int = FMM_LEAF,
                 walk_theta: float = WALK_THETA, precision: str = "fp64"):
        self._like_torch = isinstance(sources, torch.Tensor)
        self.sources = _points(sources, "sources").clone()
        self.targets = _points(targets, "targets").clone()
        m = self.targets.shape[1]
        self.exclude = (None if exclude is None
                        else _device(exclude, torch.int64).reshape(-1).clone())
        if self.exclude is not None and self.exclude.shape[0] != m:
            raise ValueError("exclude must have one entry per target")
        self.root = root
        self.order, self.theta, self.leaf_capacity = int(order), float(theta), int(leaf_capacity)
        self.walk_theta = float(walk_theta)
        self.precision = precision
        self.tree: FlatQuadtree | None = None
        self.tree_builds = 0
        self.products = 0

    @property
    def plan_builds(self) -> int:
        """Target trees and dual lists built so far (one per geometry when reused)."""
        return 0 if self.tree is None else self.tree.plan_builds

    @property
    def walks(self) -> bool:
        """electrodes.py:267-271: fewer targets than sources over FMM_AT_RATIO walk."""
        return self.targets.shape[1] * FMM_AT_RATIO < self.sources.shape[1]

    def __call__(self, charges):
        q = _device(charges).reshape(-1)
        if q.shape[0] != self.sources.shape[1]:
            raise ValueError("charges must have one entry per source")
        if self.tree is None:
            self.tree = FlatQuadtree(self.sources, q, self.root, kernel="log", order=self.order,
                                     leaf_capacity=self.leaf_capacity, precision=self.precision)
            self.tree_builds += 1
        else:
            self.tree.set_charges(q)
        self.products += 1
        if self.walks:
            phi, field = self.tree.evaluate(self.targets, exclude=self.exclude,
                                            theta=self.walk_theta)
        else:
            phi, field = self.tree.fmm_at(self.targets, exclude=self.exclude, theta=self.theta)
        like = self._like_torch or isinstance(charges, torch.Tensor)
        return _out(phi, like), _out(field, like)

    def as_linear_operator(self):
        """A SciPy LinearOperator whose matvec is the potential at the targets (square when
        the targets are as many as the sources)."""
        import numpy as np
        from scipy.sparse.linalg import LinearOperator

        m, n = self.targets.shape[1], self.sources.shape[1]

        def matvec(x):
            phi, _ = self(np.asarray(x, dtype=np.float64).reshape(-1))
            return np.asarray(phi)

        return LinearOperator((m, n), matvec=matvec, dtype=np.float64)
