"""Flat quadtree and direct sums on the GPU: the reference's flat_tree.py API, run by the
gfx950 kernels of libtd3_hip.so.

Two interaction kernels, as in the reference (flat_tree.py:1-27): ``"coulomb"``, the
softened 3D point charge ``q / sqrt(r^2 + eps^2)``, and ``"log"``, the 2D line charge
``-q ln(r^2 + eps^2)``.  Positions, targets and fields are component-major ``(2, N)``;
per-particle scalars are ``(N,)``; per-node arrays are node-major.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _abi
from ._utils import _check, _device, _is_torch, _out, _points

_KERNELS = ("coulomb", "log")
_PRECISIONS = {"fp64": _abi.FP64, "mixed": _abi.MIXED}


def _precision(precision: str) -> int:
    if precision not in _PRECISIONS:
        raise ValueError(f"precision must be one of {tuple(_PRECISIONS)}")
    return _PRECISIONS[precision]


def direct_field(
    targets,
    pos,
    charge,
    *,
    softening=0.0,
    exclude=None,
    compiled=True,
    kernel="coulomb",
    precision="fp64",
):
    """Exact potential ``(M,)`` and field ``(2, M)`` at ``targets`` ``(2, M)``, without ``k_e``.

    flat_tree.py:125-160.  ``pos`` ``(2, N)`` and ``charge`` ``(N,)`` are the sources;
    ``exclude[i]`` is a source skipped for target ``i`` (-1 for none).  ``compiled`` is
    accepted for compatibility (every call runs the GPU kernels).  ``precision="mixed"``
    evaluates logarithms and reciprocals in FP32 (sums stay FP64).
    """
    if kernel not in _KERNELS:
        raise ValueError(f"kernel must be one of {_KERNELS}")
    like_torch = _is_torch(targets) or _is_torch(pos)
    t = _points(targets, "targets")
    p = _points(pos, "pos")
    q = _device(charge).reshape(-1)
    m, n = t.shape[1], p.shape[1]
    if q.shape[0] != n:
        raise ValueError("pos and charge must have the same length")
    if exclude is None:
        ex = torch.full((m,), -1, dtype=torch.int64, device=t.device)
    else:
        ex = _device(exclude, torch.int64).reshape(-1)
    phi = torch.zeros(m, dtype=torch.float64, device=t.device)
    field = torch.zeros((2, m), dtype=torch.float64, device=t.device)
    if not m or not n:
        return _out(phi, like_torch), _out(field, like_torch)
    status = torch.zeros(1, dtype=torch.int64, device=t.device)
    nbytes = _abi.lib().td3_direct_workspace_bytes(m, n)
    ws = _abi.workspace(nbytes)
    _abi.call(
        "td3_direct",
        _abi.LOG if kernel == "log" else _abi.COULOMB,
        _precision(precision),
        _abi.ptr(t),
        m,
        _abi.ptr(ex),
        _abi.ptr(p),
        n,
        _abi.ptr(q),
        float(softening),
        _abi.ptr(phi),
        _abi.ptr(field),
        _abi.ptr(status),
        _abi.ptr(ws),
        nbytes,
        _abi.stream(),
    )
    _check(status)
    return _out(phi, like_torch), _out(field, like_torch)


#: Highest expansion order of the kernels (flat_tree.py:46, TREE_D_3D_MAX_ORDER).
MAX_ORDER = 64
#: Default orders (flat_tree.py:50-51).
LOG_ORDER = 12
COULOMB_ORDER = 4

#: Device arrays of the tree, exposed in the caller's array type.
_PUBLIC = (
    "pos",
    "charge",
    "perm",
    "node_center",
    "node_half",
    "node_start",
    "node_end",
    "node_child",
    "node_depth",
    "node_q",
    "node_qabs",
    "node_p",
    "node_s",
    "moment_re",
    "moment_im",
)

_CACHE: dict = {}


def _binomial_table(size: int) -> np.ndarray:
    """C(n, k) for 0 <= n, k < size, as floats (flat_tree.py:224-231)."""
    table = np.zeros((size, size))
    for row in range(size):
        table[row, 0] = 1.0
        for col in range(1, row + 1):
            table[row, col] = table[row - 1, col - 1] + table[row - 1, col]
    return table


def _device_table(name: str, builder, *key) -> torch.Tensor:
    """A small constant table on the current device, built once."""
    dev = _abi.device()
    k = (name, dev.index, *key)
    if k not in _CACHE:
        _CACHE[k] = torch.from_numpy(np.ascontiguousarray(builder(*key))).to(dev)
    return _CACHE[k]


class FlatNode:
    """Read-only view of one node of a `FlatQuadtree` (flat_tree.py:163-211)."""

    def __init__(self, tree: "FlatQuadtree", index: int) -> None:
        self.tree = tree
        self.index = index

    @property
    def quad(self):
        from .quad import Quad

        x, y = (float(v) for v in self.tree._d["node_center"][self.index].tolist())
        return Quad(x, y, float(self.tree._d["node_half"][self.index]))

    @property
    def count(self) -> int:
        d = self.tree._d
        return int(d["node_end"][self.index] - d["node_start"][self.index])

    @property
    def total_charge(self) -> float:
        self.tree._log_node_arrays()
        return float(self.tree._d["node_q"][self.index])

    @property
    def total_abs_charge(self) -> float:
        return float(self.tree._d["node_qabs"][self.index])

    @property
    def dipole(self) -> np.ndarray:
        self.tree._log_node_arrays()
        return self.tree._d["node_p"][self.index].cpu().numpy().copy()

    @property
    def second_moment(self) -> np.ndarray:
        self.tree._log_node_arrays()
        sxx, sxy, syy = (float(v) for v in self.tree._d["node_s"][self.index].tolist())
        return np.array([[sxx, sxy], [sxy, syy]])

    @property
    def particles(self) -> np.ndarray:
        d = self.tree._d
        start, end = int(d["node_start"][self.index]), int(d["node_end"][self.index])
        return d["perm"][start:end].cpu().numpy()

    def is_external(self) -> bool:
        return int(self.tree._d["node_child"][self.index]) < 0

    @property
    def children(self):
        first = int(self.tree._d["node_child"][self.index])
        if first < 0:
            return (None, None, None, None)
        return tuple(FlatNode(self.tree, first + k) for k in range(4))


class _Arrays(dict):
    """A tree's device arrays by name.  The cells' moments (and the cell charges that
    come with them) are made on first use, by the call `defer` was given: work queued
    after the tree that needs the cells only (the interaction rows) starts before them."""

    MOMENTS = ("node_q", "node_p", "node_s", "node_qabs", "moment_re", "moment_im")

    def __init__(self, *args, **kw):
        super().__init__(*args, **kw)
        self.pending = None

    def defer(self, make) -> None:
        for name in self.MOMENTS:
            self.pop(name, None)
        self.pending = make

    def __missing__(self, key):
        make = self.pending
        if make is None or key not in self.MOMENTS:
            raise KeyError(key)
        self.pending = None
        try:
            make()
        except BaseException:
            self.pending = make  # made on the next use instead
            raise
        return dict.__getitem__(self, key)


class _Streams:
    """A second stream (default priority) beside the caller's.  `side()` makes it
    current; from the start it runs after ``ready`` (an event of the caller's stream) or
    else after everything the caller has queued.  `mark()` and `mark_caller()` record an
    event on the second stream and on the caller's; `wait(e)` makes the caller's stream
    wait for e, `side_wait(e)` the second stream.  `used_by_caller` tells torch's allocator
    that tensors made on the second stream are read by the caller's.  On a CPU device
    (the kernels' emulation) everything stays on the one stream."""

    _side_streams: dict = {}

    def __init__(self, dev: torch.device, ready=None):
        self.on_gpu = dev.type == "cuda"
        if self.on_gpu:
            self.main = torch.cuda.current_stream(dev)
            st = _Streams._side_streams.get(dev.index)
            if st is None:
                st = _Streams._side_streams[dev.index] = torch.cuda.Stream(dev)
            self.stream = st
            if ready is not None:
                st.wait_event(ready)
            else:
                st.wait_stream(self.main)

    @staticmethod
    def record(dev: torch.device):
        if dev.type != "cuda":
            return None
        event = torch.cuda.Event()
        event.record()
        return event

    def side(self):
        import contextlib

        return torch.cuda.stream(self.stream) if self.on_gpu else contextlib.nullcontext()

    def mark(self):
        if not self.on_gpu:
            return None
        event = torch.cuda.Event()
        event.record(self.stream)
        return event

    def mark_caller(self):
        if not self.on_gpu:
            return None
        event = torch.cuda.Event()
        event.record(self.main)
        return event

    def wait(self, event) -> None:
        if event is not None:
            self.main.wait_event(event)

    def side_wait(self, event) -> None:
        if event is not None:
            self.stream.wait_event(event)

    def used_by_caller(self, *tensors) -> None:
        if self.on_gpu:
            for t in tensors:
                if t is not None:
                    t.record_stream(self.main)


class FlatQuadtree:
    """Array quadtree for point charges in a plane or for line charges, on the GPU.

    The reference's FlatQuadtree (flat_tree.py:437-1676): same constructor, same tree
    (bit-identical arrays), same ``evaluate`` / ``fmm`` / ``fmm_at`` / ``field_at`` /
    ``set_charges``.  ``precision`` selects all-FP64 (``"fp64"``, the reference's
    arithmetic) or ``"mixed"`` (FP32 far field and near field on FP64 tree data, FP64
    sums); every call may override it.
    """

    def __init__(
        self,
        positions,
        charges,
        root_quad,
        *,
        leaf_capacity: int = 8,
        max_depth: int = 48,
        compiled: bool = True,
        kernel: str = "coulomb",
        order: int | None = None,
        symmetric: bool = False,
        comm=None,
        precision: str = "fp64",
    ) -> None:
        from ._utils import _finite

        if kernel not in _KERNELS:
            raise ValueError(f"kernel must be one of {_KERNELS}")
        if order is None:
            order = LOG_ORDER if kernel == "log" else COULOMB_ORDER
        if isinstance(order, bool) or int(order) != order or order < 1:
            raise ValueError("order must be a positive integer")
        if order > MAX_ORDER:
            raise ValueError(f"order must be at most {MAX_ORDER}, got {order}")
        if comm is not None:
            raise NotImplementedError("tree-d-3d_amd runs on one GPU (comm must be None)")
        self.kernel = kernel
        self.order = int(order)
        self.symmetric = bool(symmetric)
        self.comm = None
        self.precision = precision
        _precision(precision)
        _finite("half_size", root_quad.half, positive=True)
        _finite("xmid", root_quad.xmid)
        _finite("ymid", root_quad.ymid)
        for name, value in (("leaf_capacity", leaf_capacity), ("max_depth", max_depth)):
            if isinstance(value, bool) or int(value) != value or value < 1:
                raise ValueError(f"{name} must be a positive integer")
        self._like_torch = _is_torch(positions)
        self._host_cache: dict = {}
        self._d = _Arrays()
        pos = _points(positions, "positions").clone()
        charge = _device(charges).reshape(-1).clone()
        n = pos.shape[1]
        if charge.shape[0] != n:
            raise ValueError("positions and charges must have the same length")
        self._d["pos"] = pos
        self._d["charge"] = charge
        # the positions are checked (finite, inside the closed root square: flat_tree.py:
        # 785-794) by the build's first kernel, read back with the split counts
        self.leaf_capacity = int(leaf_capacity)
        self.max_depth = int(max_depth)
        self._root = (float(root_quad.xmid), float(root_quad.ymid), float(root_quad.half))
        self._traversals: dict = {}
        self.fmm_pairs = (0, 0)
        if not n:
            self._empty()
            return
        self._build()
        self._sorted_charge()
        self._d.defer(self._moments)

    # ---------------------------------------------------------------- arrays
    def __getattr__(self, name):
        if name in _PUBLIC:
            if name in ("node_q", "node_p", "node_s") and "_d" in self.__dict__:
                self._log_node_arrays()
            try:
                t = self.__dict__["_d"][name]
            except KeyError:
                raise AttributeError(name) from None
            if self.__dict__.get("_like_torch", True):
                return t
            cache = self.__dict__["_host_cache"]
            if name not in cache:
                cache[name] = t.cpu().numpy()
            return cache[name]
        raise AttributeError(name)

    def _invalidate_host(self, *names):
        for name in names or tuple(self._host_cache):
            self._host_cache.pop(name, None)

    def _empty(self) -> None:
        """flat_tree.py:466-481: the lone root of an empty tree."""
        dev = self._d["pos"].device
        xmid, ymid, half = self._root
        p1 = self.order + 1
        self.node_count = 1
        self.levels = np.array([0, 1])
        self._d.update(
            perm=torch.empty(0, dtype=torch.int64, device=dev),
            node_center=torch.tensor([[xmid, ymid]], dtype=torch.float64, device=dev),
            node_half=torch.tensor([half], dtype=torch.float64, device=dev),
            node_start=torch.zeros(1, dtype=torch.int64, device=dev),
            node_end=torch.zeros(1, dtype=torch.int64, device=dev),
            node_child=torch.full((1,), -1, dtype=torch.int64, device=dev),
            node_depth=torch.zeros(1, dtype=torch.int64, device=dev),
            node_q=torch.zeros(1, dtype=torch.float64, device=dev),
            node_qabs=torch.zeros(1, dtype=torch.float64, device=dev),
            node_p=torch.zeros((1, 2), dtype=torch.float64, device=dev),
            node_s=torch.zeros((1, 3), dtype=torch.float64, device=dev),
            moment_re=torch.zeros((1, p1), dtype=torch.float64, device=dev),
            moment_im=torch.zeros((1, p1), dtype=torch.float64, device=dev),
        )

    def _build(self) -> None:
        """flat_tree.py:532-625 / :483-530: one td3_tree_build call, regrown on overflow."""
        pos = self._d["pos"]
        dev = pos.device
        n = pos.shape[1]
        capacity = 4 * (2 * n // self.leaf_capacity + 16)
        xmid, ymid, half = self._root
        out = np.zeros(3, dtype=np.int64)
        levels_host = np.zeros(67, dtype=np.int64)
        while True:
            f64 = dict(dtype=torch.float64, device=dev)
            i64 = dict(dtype=torch.int64, device=dev)
            center = torch.empty((capacity, 2), **f64)
            half_arr = torch.empty(capacity, **f64)
            start = torch.empty(capacity, **i64)
            end = torch.empty(capacity, **i64)
            child = torch.empty(capacity, **i64)
            depth = torch.empty(capacity, **i64)
            levels = torch.zeros(67, **i64)
            perm = torch.empty(n, **i64)
            sorted_pos = torch.empty((2, n), **f64)
            info = torch.empty((capacity, 4), dtype=torch.int32, device=dev)
            nbytes = _abi.lib().td3_tree_workspace_bytes(n, capacity)
            ws = _abi.workspace(nbytes)
            code = _abi.lib().td3_tree_build(
                _abi.ptr(pos), n, xmid, ymid, half, self.leaf_capacity, self.max_depth, capacity,
                _abi.ptr(center), _abi.ptr(half_arr), _abi.ptr(start), _abi.ptr(end),
                _abi.ptr(child), _abi.ptr(depth), _abi.ptr(levels), _abi.ptr(perm),
                _abi.ptr(sorted_pos), _abi.ptr(info), _abi.ptr(ws), nbytes,
                out.ctypes.data, levels_host.ctypes.data, _abi.stream(),
            )
            if code == _abi.TD3_ERR_CAPACITY:
                capacity = max(2 * capacity, int(out[2]))
                continue
            if code == _abi.TD3_ERR_ARG:
                message = _abi.last_error()
                for known in ("Particle position must be finite",
                              "Particle lies outside tree bounds"):
                    if known in message:
                        raise ValueError(known)
            _abi.check(code, "td3_tree_build")
            break
        count, n_levels = int(out[0]), int(out[1])
        self.node_count = count
        self.levels = levels_host[: n_levels + 1].copy()
        self._d.update(
            perm=perm,
            node_center=center[:count],
            node_half=half_arr[:count],
            node_start=start[:count],
            node_end=end[:count],
            node_child=child[:count],
            node_depth=depth[:count],
        )
        self._sorted_pos = sorted_pos
        self._node_info = info[:count]
        self._depth = n_levels - 1

    def _sorted_charge(self) -> None:
        n = self._d["pos"].shape[1]
        out = torch.empty(n, dtype=torch.float64, device=self._d["pos"].device)
        _abi.call("td3_permute", _abi.ptr(self._d["charge"]), 1, n, _abi.ptr(self._d["perm"]), n,
                  _abi.ptr(out), _abi.stream())
        self._scharge = out
        # work on a second stream that needs the cells and the sorted charges starts here
        self._ready = _Streams.record(self._d["pos"].device)

    def _moments(self) -> None:
        """flat_tree.py:663-744: the cells' charges and multipole moments."""
        dev = self._d["pos"].device
        c, n = self.node_count, self._d["pos"].shape[1]
        f64 = dict(dtype=torch.float64, device=dev)
        # written in full by the kernels (log: q and p from the moments; s stays zero)
        # log: q and p are read off the moments only when asked for (_log_node_arrays)
        node_qabs = torch.empty(c, **f64)
        if self.kernel == "log":
            node_q = node_p = node_s = None
        else:
            node_q = torch.empty(c, **f64)
            node_p = torch.empty((c, 2), **f64)
            node_s = torch.empty((c, 3), **f64)
        p1 = self.order + 1
        if self.kernel == "log":
            mre = torch.empty((c, p1), **f64)
            mim = torch.empty((c, p1), **f64)
        else:
            mre = torch.zeros((1, p1), **f64)
            mim = torch.zeros((1, p1), **f64)
        bsz = 2 * self.order + 1
        binom = _device_table("binomial", _binomial_table, bsz)
        levels = np.ascontiguousarray(self.levels, dtype=np.int64)
        sdip = self.__dict__.get("_sdip")
        if sdip is not None:
            _abi.call(
                "td3_log_moments_dipole", self.order, _abi.ptr(self._sorted_pos),
                _abi.ptr(self._scharge), _abi.ptr(sdip), n, _abi.ptr(self._d["node_center"]),
                _abi.ptr(self._node_info), c, levels.ctypes.data, len(levels) - 1,
                _abi.ptr(binom), bsz, _abi.ptr(node_qabs), _abi.ptr(mre), _abi.ptr(mim),
                _abi.stream(),
            )
        else:
            _abi.call(
                "td3_moments", _abi.LOG if self.kernel == "log" else _abi.COULOMB, self.order,
                _abi.ptr(self._sorted_pos), _abi.ptr(self._scharge), n,
                _abi.ptr(self._d["node_center"]), _abi.ptr(self._node_info), c, levels.ctypes.data,
                len(levels) - 1, _abi.ptr(binom), bsz, _abi.ptr(node_q), _abi.ptr(node_qabs),
                _abi.ptr(node_p), _abi.ptr(node_s), _abi.ptr(mre), _abi.ptr(mim), _abi.stream(),
            )
        for name in ("node_q", "node_p", "node_s"):
            self._d.pop(name, None)
        if self.kernel != "log":
            self._d.update(node_q=node_q, node_p=node_p, node_s=node_s)
        self._d.update(node_qabs=node_qabs, moment_re=mre, moment_im=mim)
        self._invalidate_host("node_q", "node_qabs", "node_p", "node_s", "moment_re",
                              "moment_im")

    def _log_node_arrays(self) -> None:
        """The log tree's node_q, node_p (a_0 and a_1 of the moments) and node_s (zero),
        made when first asked for: the FMM itself reads the moments."""
        d = self._d
        if self.kernel != "log" or "node_q" in d:
            return
        mre, mim = d["moment_re"], d["moment_im"]
        d["node_q"] = mre[:, 0].clone()
        d["node_p"] = torch.stack((mre[:, 1], mim[:, 1]), dim=1).contiguous()
        d["node_s"] = torch.zeros((mre.shape[0], 3), dtype=torch.float64, device=mre.device)
        self._invalidate_host("node_q", "node_p", "node_s")

    def set_charges(self, charges, dipoles=None) -> None:
        """flat_tree.py:853-868: new charges for the same particles; keeps cells and lists.

        ``dipoles`` ``(2, N)`` (log kernel only; no counterpart in the reference) gives the
        particles dipoles ``D`` as well, each adding ``-2 D.d / (r^2 + eps^2)`` to the
        potential at ``d = target - particle`` (``-D`` times the gradient of a unit
        charge's potential in its position) and minus its gradient to the field.  `fmm`
        and `fmm_at` include them; a new `set_charges` without dipoles drops them.
        """
        charges = _device(charges).reshape(-1)
        n = self._d["pos"].shape[1]
        if charges.shape[0] != n:
            raise ValueError("charges must have one entry per particle")
        dip = None
        if dipoles is not None:
            if self.kernel != "log":
                raise ValueError("dipoles need the log kernel")
            dip = _device(dipoles)
            if dip.shape != (2, n):
                raise ValueError("dipoles must have shape (2, N) (component-major)")
            if not bool(torch.isfinite(dip).all()):
                raise ValueError("dipoles must be finite")
        self._d["charge"].copy_(charges)
        self._invalidate_host("charge")
        self._cart = None
        self._cart_f32 = None
        self._scaled = {}
        self._src4 = None
        self._sdip = None
        if n:
            self._sorted_charge()
            if dip is not None:
                sdip = torch.empty((2, n), dtype=torch.float64, device=dip.device)
                _abi.call("td3_permute", _abi.ptr(dip.contiguous()), 2, n,
                          _abi.ptr(self._d["perm"]), n, _abi.ptr(sdip), _abi.stream())
                self._sdip = sdip
            self._d.defer(self._moments)

    @property
    def root(self) -> FlatNode:
        return FlatNode(self, 0)

    def collect_nodes(self) -> list:
        """Views of every node that holds at least one particle."""
        occupied = torch.nonzero(self._d["node_end"] > self._d["node_start"]).reshape(-1)
        return [FlatNode(self, int(i)) for i in occupied.cpu().tolist()]

    # ---------------------------------------------------------------- the walk
    def _target_order(self, targets: torch.Tensor) -> torch.Tensor:
        """flat_tree.py:891-912: tree order for the sources themselves, else Morton order."""
        pos = self._d["pos"]
        if targets.shape == pos.shape and bool(torch.equal(targets, pos)):
            return self._d["perm"]
        m = targets.shape[1]
        lo, hi = torch.aminmax(targets, dim=1)
        span = float((hi - lo).max())
        scale = (65535.0 / span) if span > 0 else 0.0
        order = torch.empty(m, dtype=torch.int64, device=targets.device)
        nbytes = _abi.lib().td3_target_order_workspace_bytes(m)
        ws = _abi.workspace(nbytes)
        lox, loy = (float(v) for v in lo.tolist())
        _abi.call("td3_target_order", _abi.ptr(targets), m, lox, loy, scale, _abi.ptr(order),
                  _abi.ptr(ws), nbytes, _abi.stream())
        return order

    def evaluate(self, targets, *, exclude=None, theta=0.5, softening=0.0, potential=True,
                 precision=None):
        """Potential ``(M,)`` and field ``(2, M)`` at ``targets`` ``(2, M)``, without ``k_e``.

        flat_tree.py:988-1061: the Barnes-Hut walk.  ``exclude[i]`` is a source index
        skipped for target ``i`` (-1: none); ``potential=False`` returns the potential as
        zeros (and skips the line-charge logarithms).
        """
        if self.__dict__.get("_sdip") is not None:
            raise ValueError("the walk takes charges only; dipoles are used by fmm and fmm_at")
        like_torch = _is_torch(targets) or self._like_torch
        t = _points(targets, "targets")
        m = t.shape[1]
        if exclude is None:
            ex = torch.full((m,), -1, dtype=torch.int64, device=t.device)
        else:
            ex = _device(exclude, torch.int64).reshape(-1)
        phi = torch.zeros(m, dtype=torch.float64, device=t.device)
        field = torch.zeros((2, m), dtype=torch.float64, device=t.device)
        n = self._d["pos"].shape[1]
        if n and m:
            status = torch.zeros(1, dtype=torch.int64, device=t.device)
            order = self._target_order(t)
            self._log_node_arrays()
            d = self._d
            _abi.call(
                "td3_walk", _abi.LOG if self.kernel == "log" else _abi.COULOMB,
                _precision(precision or self.precision), _abi.ptr(t), m, _abi.ptr(order),
                _abi.ptr(ex), _abi.ptr(self._sorted_pos), _abi.ptr(self._scharge),
                _abi.ptr(d["perm"]), n, _abi.ptr(d["node_center"]), _abi.ptr(d["node_half"]),
                _abi.ptr(self._node_info), _abi.ptr(d["node_qabs"]), _abi.ptr(d["node_q"]),
                _abi.ptr(d["node_p"]), _abi.ptr(d["node_s"]), _abi.ptr(d["moment_re"]),
                _abi.ptr(d["moment_im"]), self.order, self._depth, None, float(theta),
                float(softening), 1 if potential else 0, _abi.ptr(phi), _abi.ptr(field),
                _abi.ptr(status), _abi.stream(),
            )
            _check(status)
        if not potential:
            phi.zero_()
        return _out(phi, like_torch), _out(field, like_torch)

    def field_at(self, positions, *, theta=0.5, k_e=None, softening=0.0, reference_radius=1.0):
        """Potential ``(n,)`` and field ``(2, n)`` at probes, with ``k_e`` (flat_tree.py:1648-1676).

        For line charges the potential is ``-2 k_e sum q ln(r / reference_radius)``.
        """
        from ._utils import _finite
        from .constants import K_E

        k_e = K_E if k_e is None else k_e
        theta = _finite("theta", theta, minimum=0)
        reference_radius = _finite("reference_radius", reference_radius, positive=True)
        k_e = _finite("k_e", k_e, minimum=0)
        softening = _finite("softening", softening, minimum=0)
        like_torch = _is_torch(positions) or self._like_torch
        p = _points(positions, "positions")
        if not bool(torch.isfinite(p).all()):
            raise ValueError("positions must be finite")
        if k_e == 0:
            m = p.shape[1]
            z = torch.zeros(m, dtype=torch.float64, device=p.device)
            return _out(z, like_torch), _out(torch.zeros((2, m), dtype=torch.float64,
                                                         device=p.device), like_torch)
        phi, field = self.evaluate(p, theta=theta, softening=softening)
        if self.kernel == "log":
            phi = phi + 2 * float(self._d["charge"].sum()) * float(np.log(reference_radius))
        return _out(k_e * phi, like_torch), _out(k_e * field, like_torch)

    # ---------------------------------------------------------------- the FMM
    def _m2l_coefficients(self) -> torch.Tensor:
        """The real M2L matrix coeff[k, l] (flat_tree.py:234-245), on the device."""
        return _device_table("m2l", _m2l_coefficients, self.order, self.symmetric)

    def _lists(self, theta: float):
        """The M2L and P2P lists of the traversal as rows by node, kept per theta.

        flat_tree.py:1318-1358 (the lists depend on the cells only, so `set_charges`
        keeps them).
        """
        key = ("lists", theta)
        if key in self._traversals:
            return self._traversals[key]
        lists = self._rows(theta)
        self._traversals[key] = lists
        return lists

    def _rows(self, theta: float, deferred: bool = False):
        """td3_rows: every node's rows from the pairs it inherits, level by level.

        ``deferred``: no read-back; returns the rows at full capacity and their counts on
        the device ``(m2l, p2p, leaves, overflow)`` for `_accept_rows` to check later (an
        overflow leaves every row empty)."""
        dev = self._d["pos"].device
        c = self.node_count
        levels = np.ascontiguousarray(self.levels, dtype=np.int64)
        n_levels = len(levels) - 1
        width = self.__dict__.get("_row_width",
                                  max(self._ROW_WIDTH, FlatQuadtree._row_width_seen))
        caps = [f * c + 64 for f in self._ROW_CAPS] + [width]
        out = np.zeros(4, dtype=np.int64)
        if deferred:
            i32 = dict(dtype=torch.int32, device=dev)
            mrow, prow = torch.empty(c + 1, **i32), torch.empty(c + 1, **i32)
            # no destination per entry: the M2L walks the rows by their starts (td3_rows
            # leaves m2l_dst alone when it is NULL; _row_dst makes it if needed later)
            mcol, mdst = torch.empty(caps[0], **i32), None
            pcol = torch.empty(caps[1], **i32)
            dcounts = torch.zeros(4, dtype=torch.int64, device=dev)
            nbytes = _abi.lib().td3_rows_workspace_bytes(levels.ctypes.data, n_levels, caps[2])
            ws = _abi.workspace(nbytes)
            _abi.call("td3_rows", _abi.ptr(self._d["node_center"]), _abi.ptr(self._d["node_half"]),
                      _abi.ptr(self._node_info), levels.ctypes.data, n_levels, float(theta),
                      caps[0], caps[1], caps[2], _abi.ptr(mrow), _abi.ptr(mcol), _abi.ptr(mdst),
                      _abi.ptr(prow), _abi.ptr(pcol), _abi.ptr(ws), nbytes, None,
                      _abi.ptr(dcounts), _abi.stream())
            return (mrow, mcol, mdst), (prow, pcol, None), dcounts
        while True:
            i32 = dict(dtype=torch.int32, device=dev)
            mrow, prow = torch.empty(c + 1, **i32), torch.empty(c + 1, **i32)
            mcol, mdst = torch.empty(caps[0], **i32), torch.empty(caps[0], **i32)
            pcol = torch.empty(caps[1], **i32)
            nbytes = _abi.lib().td3_rows_workspace_bytes(levels.ctypes.data, n_levels, caps[2])
            ws = _abi.workspace(nbytes)
            code = _abi.lib().td3_rows(
                _abi.ptr(self._d["node_center"]), _abi.ptr(self._d["node_half"]),
                _abi.ptr(self._node_info), levels.ctypes.data, n_levels, float(theta), caps[0],
                caps[1], caps[2], _abi.ptr(mrow), _abi.ptr(mcol), _abi.ptr(mdst),
                _abi.ptr(prow), _abi.ptr(pcol), _abi.ptr(ws), nbytes, out.ctypes.data, None,
                _abi.stream(),
            )
            if code == _abi.TD3_ERR_CAPACITY:
                # out[3] > 1: a node's slots overflowed (the width it needs); the rows'
                # totals grow either way
                caps = [max(2 * caps[0], int(out[0]) + 64), max(2 * caps[1], int(out[1]) + 64),
                        max(caps[2], int(out[3]) + 8) if out[3] > 1 else caps[2]]
                self._row_width = caps[2]
                # later trees of the process start at the width this one needed: a time
                # step's next tree is much like this one, and a too-narrow first build
                # costs a second build (and, without a read-back, a second FMM)
                FlatQuadtree._row_width_seen = max(FlatQuadtree._row_width_seen, caps[2])
                continue
            _abi.check(code, "td3_rows")
            break
        return self._sliced_rows((mrow, mcol, mdst), (prow, pcol, None), out)

    @staticmethod
    def _sliced_rows(m2l_rows, p2p_rows, out):
        n_m2l, n_p2p, leaves = int(out[0]), int(out[1]), int(out[2])
        counts = (n_m2l // 2, (n_p2p + leaves) // 2)
        (mrow, mcol, mdst), (prow, pcol, _) = m2l_rows, p2p_rows
        if mdst is not None:
            mdst = mdst[: max(n_m2l, 1)]
        return (mrow, mcol[: max(n_m2l, 1)], mdst), \
            (prow, pcol[: max(n_p2p, 1)], None), counts

    def _traverse_lists(self, theta: float):
        """td3_traverse (pair lists by generation) + td3_lists (rows, each sorted)."""
        dev = self._d["pos"].device
        c = self.node_count
        caps = [48 * c + 64, 16 * c + 64, 10 * c + 64]
        out = np.zeros(4, dtype=np.int64)
        while True:
            m2l = torch.empty((caps[0], 2), dtype=torch.int32, device=dev)
            p2p = torch.empty((caps[1], 2), dtype=torch.int32, device=dev)
            nbytes = _abi.lib().td3_traverse_workspace_bytes(caps[2])
            ws = _abi.workspace(nbytes)
            code = _abi.lib().td3_traverse(
                _abi.ptr(self._d["node_center"]), _abi.ptr(self._d["node_half"]),
                _abi.ptr(self._node_info), c, self._depth, float(theta), caps[0], caps[1],
                caps[2], _abi.ptr(m2l), _abi.ptr(p2p), _abi.ptr(ws), nbytes, out.ctypes.data,
                _abi.stream(),
            )
            if code == _abi.TD3_ERR_CAPACITY:
                caps = [max(2 * caps[i], int(out[i]) + 64) for i in range(3)]
                continue
            _abi.check(code, "td3_traverse")
            break
        n_m2l, n_p2p = int(out[0]), int(out[1])
        rows = []
        for pairs, count, self_once in ((m2l, n_m2l, 0), (p2p, n_p2p, 1)):
            row = torch.empty(c + 1, dtype=torch.int32, device=dev)
            col = torch.empty(max(2 * count, 1), dtype=torch.int32, device=dev)
            dst = torch.empty(max(2 * count, 1), dtype=torch.int32, device=dev)
            nbytes = _abi.lib().td3_lists_workspace_bytes(c, count, self_once)
            ws = _abi.workspace(nbytes)
            _abi.call("td3_lists", _abi.ptr(pairs), count, self_once, c, c, _abi.ptr(row),
                      _abi.ptr(col), _abi.ptr(dst), _abi.ptr(ws), nbytes, _abi.stream())
            rows.append((row, col, dst))
        return (rows[0], rows[1], (n_m2l, n_p2p))

    def _sources4(self) -> torch.Tensor:
        """(N, 4) float64 (x, y, q, 0) of the sorted sources (kept until `set_charges`)."""
        cached = self.__dict__.get("_src4")
        if cached is not None:
            return cached
        n = self._d["pos"].shape[1]
        src4 = (torch.empty if n else torch.zeros)((max(n, 1), 4), dtype=torch.float64,
                                                    device=self._d["pos"].device)
        _abi.call("td3_pack_sources", _abi.ptr(self._sorted_pos), _abi.ptr(self._scharge), n,
                  _abi.ptr(src4), _abi.stream())
        self._src4 = src4
        return src4

    def _leaf_index(self) -> torch.Tensor:
        """(N,) int32: the leaf of every sorted particle (kept with the tree)."""
        cached = self.__dict__.get("_leaf_of")
        if cached is not None:
            return cached
        n = self._d["pos"].shape[1]
        leaf_of = (torch.empty if n else torch.zeros)(max(n, 1), dtype=torch.int32,
                                                       device=self._d["pos"].device)
        _abi.call("td3_leaf_index", _abi.ptr(self._node_info), self.node_count,
                  _abi.ptr(leaf_of), _abi.stream())
        self._leaf_of = leaf_of
        return leaf_of

    def _scaled_moments(self, precision: int) -> tuple:
        key = ("scaled", precision)
        cached = self.__dict__.get("_scaled", {})
        if key in cached:
            return cached[key]
        c, p1 = self.node_count, self.order + 1
        dtype = torch.float64 if precision == _abi.FP64 else torch.float32
        dev = self._d["pos"].device
        re = torch.empty((c, p1), dtype=dtype, device=dev)
        im = torch.empty((c, p1), dtype=dtype, device=dev)
        _abi.call("td3_scale_moments", precision, _abi.ptr(self._d["moment_re"]),
                  _abi.ptr(self._d["moment_im"]), _abi.ptr(self._d["node_half"]), c, self.order,
                  _abi.ptr(re), _abi.ptr(im), _abi.stream())
        cached[key] = (re, im)
        self._scaled = cached
        return re, im

    def _m2l_records(self) -> torch.Tensor:
        """The FP32 M2L's source records of every cell (td3_m2l_records; order <= 15),
        kept with the scaled moments until `set_charges`."""
        cached = self.__dict__.get("_scaled", {})
        if "records" in cached:
            return cached["records"]
        nbytes = _abi.lib().td3_m2l_record_bytes(self.order)
        rec = torch.empty(max(self.node_count * nbytes // 4, 1), dtype=torch.float32,
                          device=self._d["pos"].device)
        _abi.call("td3_m2l_records", self.order, _abi.ptr(self._d["moment_re"]),
                  _abi.ptr(self._d["moment_im"]), _abi.ptr(self._d["node_center"]),
                  _abi.ptr(self._d["node_half"]), self.node_count, _abi.ptr(rec), _abi.stream())
        cached["records"] = rec
        self._scaled = cached
        return rec

    #: target trees and dual lists `fmm_at` has built (kept per targets, theta and leaf)
    plan_builds = 0

    #: in mixed precision, M2L into cells at depth <= this stays FP64 (their terms are the
    #: largest: a cell's net charge grows with its size)
    FP64_TOP_DEPTH = 5

    #: share of the targets (the first in tree order) whose near field the FMM sums as soon
    #: as the rows exist, beside the moments' last levels; the rest wait for the M2L and run
    #: beside the L2L (measured at N = 1M: 0.2 gives 2.30 ms a step against 2.32 ms at 0)
    NEAR_EARLY = 0.2

    #: first capacities of the interaction rows per node (M2L and P2P entries in all):
    #: exceeded, the rows are built again with twice the room
    _ROW_CAPS = (96, 40)
    #: entries per node of the row construction's slots (M2L row, P2P row, hand-downs to
    #: the children); a node that needs more has the rows built again with its width
    _ROW_WIDTH = 128
    #: the widest slots a tree of this process has needed (0: none wider than _ROW_WIDTH)
    _row_width_seen = 0

    def _local_expansions(self, theta: float, precision: int, target=None, m2l_rows=None):
        """Scaled local expansions B_l = b_l r^l of every cell of ``target`` (this tree by
        default): M2L from this tree's cells along ``m2l_rows``, then L2L down ``target``."""
        target = self if target is None else target
        if m2l_rows is None:
            m2l_rows = self._lists(theta)[0]
        lre, lim = self._m2l(precision, target, m2l_rows)
        self._l2l(target, lre, lim)
        return lre, lim

    def _m2l(self, precision: int, target, m2l_rows):
        """Every cell's scaled local expansion from the M2L along its row (zeros without
        pairs), before L2L."""
        mrow, mcol, mdst = m2l_rows
        c, p1 = target.node_count, self.order + 1
        dev = self._d["pos"].device
        # every row is written by the M2L of its range (zeros without pairs)
        lre = torch.empty((c, p1), dtype=torch.float64, device=dev)
        lim = torch.empty((c, p1), dtype=torch.float64, device=dev)
        coeff = self._m2l_coefficients()
        levels = np.ascontiguousarray(target.levels, dtype=np.int64)
        if precision == _abi.FP64:
            ranges = [(_abi.FP64, 0, c)]
        else:
            split = int(levels[min(self.FP64_TOP_DEPTH + 1, len(levels) - 1)])
            ranges = [(_abi.FP64, 0, split), (_abi.MIXED, split, c)]
        td = target._d
        for prec, lo, hi in ranges:
            if hi <= lo:
                continue
            if prec != _abi.FP64 and self.order + 1 <= 16:
                # FP32 on the matrix cores from the source records: no workspace, no pair
                # count needed
                _abi.call("td3_m2l_f32", self.order, _abi.ptr(mcol), _abi.ptr(mrow), lo, hi,
                          _abi.ptr(td["node_center"]), _abi.ptr(td["node_half"]),
                          _abi.ptr(self._m2l_records()), _abi.ptr(coeff), _abi.ptr(lre),
                          _abi.ptr(lim), _abi.stream())
                continue
            are, aim = self._scaled_moments(prec)
            if precision != _abi.FP64 and self.order + 1 <= 16:
                # the FP64 top levels of the mixed run on the matrix cores, gridded by rows
                # (no workspace, no pair count needed)
                nbytes, ws = 0, None
            else:
                nbytes = _abi.lib().td3_m2l_workspace_bytes(int(mcol.shape[0]), self.order,
                                                            prec)
                ws = _abi.workspace(nbytes)
                if mdst is None:
                    mdst = self._row_dst(mrow, mcol)
            _abi.call("td3_m2l", prec, self.order, _abi.ptr(mcol), _abi.ptr(mdst),
                      _abi.ptr(mrow), lo, hi, _abi.ptr(td["node_center"]),
                      _abi.ptr(td["node_half"]), _abi.ptr(self._d["node_center"]),
                      _abi.ptr(self._d["node_half"]), _abi.ptr(are), _abi.ptr(aim),
                      _abi.ptr(coeff), _abi.ptr(lre), _abi.ptr(lim), _abi.ptr(ws), nbytes,
                      _abi.stream())
        return lre, lim

    def _row_dst(self, mrow, mcol):
        """Each M2L entry's row (its destination cell), made from the row starts when the
        rows were built without it (the mixed FMM's rows: only the lane-per-pair M2L reads
        it), and kept with the rows."""
        cache = self.__dict__.setdefault("_row_dst_cache", {})
        hit = cache.get(mcol.data_ptr())
        if hit is not None and hit[0] is mcol:
            return hit[1]
        total = int(mrow[-1].item())
        dev = mrow.device
        if total == 0:
            dst = torch.zeros(1, dtype=torch.int32, device=dev)
        else:
            counts = (mrow[1:] - mrow[:-1]).to(torch.int64)
            dst = torch.repeat_interleave(
                torch.arange(mrow.shape[0] - 1, dtype=torch.int32, device=dev), counts,
                output_size=total)
        cache[mcol.data_ptr()] = (mcol, dst)
        return dst

    def _l2l(self, target, lre, lim) -> None:
        """L2L down ``target``, level by level, in place."""
        levels = np.ascontiguousarray(target.levels, dtype=np.int64)
        bsz = 2 * self.order + 1
        binom = _device_table("binomial", _binomial_table, bsz)
        _abi.call("td3_l2l", self.order, _abi.ptr(target._d["node_center"]),
                  _abi.ptr(target._d["node_half"]), _abi.ptr(target._node_info),
                  levels.ctypes.data, len(levels) - 1, _abi.ptr(binom), bsz, _abi.ptr(lre),
                  _abi.ptr(lim), _abi.stream())

    def local_expansions(self, theta: float = 0.5, precision=None):
        """The reference's (unscaled) local expansions ``local_re``, ``local_im`` after L2L."""
        lre, lim = self._local_expansions(float(theta), _precision(precision or self.precision))
        _abi.call("td3_unscale_locals", _abi.ptr(self._d["node_half"]), self.node_count,
                  self.order, _abi.ptr(lre), _abi.ptr(lim), _abi.stream())
        return _out(lre, self._like_torch), _out(lim, self._like_torch)

    @staticmethod
    def _fmm_out(phi, field, jac, like_torch):
        if jac is None:
            return _out(phi, like_torch), _out(field, like_torch)
        return _out(phi, like_torch), _out(field, like_torch), _out(jac, like_torch)

    def fmm(self, *, theta: float = 0.5, softening: float = 0.0, potential: bool = True,
            precision=None, jacobian: bool = False):
        """Potential ``(N,)`` and field ``(2, N)`` at every source by the FMM, without ``k_e``.

        flat_tree.py:1461-1634: dual-tree traversal, M2L both ways, L2L, L2P and exact P2P
        between near leaves (softened; the far field of the log kernel is unsoftened).
        Each particle's own charge is excluded.  ``potential=False`` returns zeros for the
        potential and skips the near-field logarithms.

        Log kernel only, no counterpart in the reference: ``jacobian=True`` also returns
        ``dE/dx`` at every source, ``(3, N)`` rows xx, xy, yy; dipoles given to
        `set_charges` are included.  Both evaluate the near field in FP64.
        """
        from ._utils import _finite

        theta = _finite("theta", theta, minimum=0)
        if not 0 < theta < 1:
            raise ValueError("FMM theta must lie in (0, 1)")
        softening = _finite("softening", softening, minimum=0)
        n = self._d["pos"].shape[1]
        dev = self._d["pos"].device
        sdip = self.__dict__.get("_sdip")
        _check_jacobian(self, jacobian, sdip)
        # every target's values are written by the evaluation
        alloc = torch.empty if n else torch.zeros
        phi = alloc(n, dtype=torch.float64, device=dev)
        field = alloc((2, n), dtype=torch.float64, device=dev)
        jac = alloc((3, n), dtype=torch.float64, device=dev) if jacobian else None
        if not n:
            return self._fmm_out(phi, field, jac, self._like_torch)
        if self.kernel != "log":
            return self._coulomb_fmm(theta, softening, potential, precision)
        prec = _precision(precision or self.precision)
        if not jacobian and sdip is None:
            return self._fmm_overlapped(theta, softening, potential, prec, phi, field)
        # the adjoint's evaluation (dipole sources, the field's Jacobian) takes rows that were
        # read back
        lre, lim = self._local_expansions(theta, prec)
        _, (prow, pcol, _), counts = self._lists(theta)
        status = torch.zeros(1, dtype=torch.int64, device=dev)
        _abi.call(
            "td3_log_fmm_eval_ext", self.order, _abi.ptr(self._sorted_pos),
            _abi.ptr(self._d["perm"]), n, _abi.ptr(self._d["node_center"]),
            _abi.ptr(self._d["node_half"]), _abi.ptr(self._node_info), _abi.ptr(lre),
            _abi.ptr(lim), _abi.ptr(prow), _abi.ptr(pcol), _abi.ptr(self._leaf_index()),
            _abi.ptr(self._sources4()), _abi.ptr(sdip), softening, _abi.ptr(phi),
            _abi.ptr(field), _abi.ptr(jac), _abi.ptr(status), _abi.stream(),
        )
        _check(status)
        if not potential:
            phi.zero_()
        self.fmm_pairs = counts
        return self._fmm_out(phi, field, jac, self._like_torch)

    def _fmm_overlapped(self, theta, softening, potential, prec, phi, field):
        """The log FMM on two streams.  A second stream, from the moment the tree and the
        sorted charges exist: the leaf map, the packed sources, the interaction rows and the
        near field (td3_fmm_near) of the first NEAR_EARLY of the targets, then (once the M2L
        is done) that of the rest.  The caller's stream:
        the cells' moments and their scaled forms (beside the rows: the rows depend on the
        cells only), the M2L once the rows are there, then the L2L (beside the near field:
        its small levels leave most of the GPU to it) and the L2P with the near sums added
        (td3_fmm_far).  Every kernel computes what it computes in sequence; the results
        are the same bit for bit.

        In mixed precision (orders up to 15) the rows are built at full capacity with their
        counts left on the device and checked with the status at the end (an overflow
        leaves every row empty: the rows are built again with room to spare and the FMM
        runs again on them)."""
        n = self._d["pos"].shape[1]
        dev = self._d["pos"].device
        key = ("lists", theta)
        cached = self._traversals.get(key)
        deferred = cached is None and prec != _abi.FP64 and self.order + 1 <= 16
        streams = _Streams(dev, self.__dict__.get("_ready"))
        with streams.side():
            leaf_of, src4 = self._leaf_index(), self._sources4()
            dcounts = None
            if cached is not None:
                m2l_rows, p2p_rows, counts = cached
            elif deferred:
                m2l_rows, p2p_rows, dcounts = self._rows(theta, deferred=True)
            else:
                m2l_rows, p2p_rows, counts = self._lists(theta)
            rows_done = streams.mark()
            near = torch.empty((3, n), dtype=torch.float64, device=dev)
            status = torch.zeros(1, dtype=torch.int64, device=dev)

            def near_field(first, count):
                _abi.call("td3_fmm_near", _abi.LOG, prec, _abi.ptr(self._sorted_pos),
                          _abi.ptr(self._d["perm"]), n, first, count,
                          _abi.ptr(self._node_info), _abi.ptr(p2p_rows[0]),
                          _abi.ptr(p2p_rows[1]), _abi.ptr(leaf_of), _abi.ptr(src4), softening,
                          1 if potential else 0, _abi.ptr(near), _abi.ptr(status),
                          _abi.stream())

            # the first targets' near field at once, beside the moments' last small levels
            early = int(n * self.NEAR_EARLY)
            near_field(0, early)
        streams.used_by_caller(leaf_of, src4, *m2l_rows, *p2p_rows, dcounts, near, status)
        # the M2L's inputs that need the moments only
        self._scaled_moments(_abi.FP64)
        if prec != _abi.FP64:
            if self.order + 1 <= 16:
                self._m2l_records()
            else:
                self._scaled_moments(prec)
        streams.wait(rows_done)
        lre, lim = self._m2l(prec, self, m2l_rows)
        m2l_done = streams.mark_caller()
        with streams.side():
            streams.side_wait(m2l_done)
            near_field(early, n - early)
            near_done = streams.mark()
        self._l2l(self, lre, lim)
        streams.wait(near_done)
        _abi.call("td3_fmm_far", _abi.LOG, self.order, _abi.ptr(self._sorted_pos),
                  _abi.ptr(self._d["perm"]), n, _abi.ptr(self._d["node_center"]),
                  _abi.ptr(self._d["node_half"]), _abi.ptr(lre), _abi.ptr(lim),
                  _abi.ptr(leaf_of), _abi.ptr(near), 1 if potential else 0, _abi.ptr(phi),
                  _abi.ptr(field), _abi.stream())
        if deferred:
            host = torch.cat((status, dcounts)).cpu().numpy()
            if host[4]:
                # the rows overflowed (every row was left empty): build them again with
                # room to spare (read back, capacities doubled until they fit), then
                # evaluate again on them
                self._lists(theta)
                return self._fmm_overlapped(theta, softening, potential, prec, phi, field)
            lists = self._sliced_rows(m2l_rows, p2p_rows, host[1:])
            self._traversals[key] = lists
            counts = lists[2]
            _check(host[:1])
        else:
            _check(status)
        self.fmm_pairs = counts
        return _out(phi, self._like_torch), _out(field, self._like_torch)

    # ---------------------------------------------------------------- Coulomb FMM
    def _cart_moments(self):
        """Cartesian moments m_g and primed moments m'_g of every cell (kept until
        `set_charges`; flat_tree.py:1063-1084)."""
        cached = self.__dict__.get("_cart")
        if cached is not None:
            return cached
        dev = self._d["pos"].device
        nc = (self.order + 1) * (self.order + 2) // 2
        mom = torch.empty((self.node_count, nc), dtype=torch.float64, device=dev)
        prime = torch.empty_like(mom)
        bsz = 2 * self.order + 1
        binom = _device_table("binomial", _binomial_table, bsz)
        levels = np.ascontiguousarray(self.levels, dtype=np.int64)
        _abi.call("td3_cart_moments", self.order, _abi.ptr(self._sorted_pos),
                  _abi.ptr(self._scharge), self._d["pos"].shape[1],
                  _abi.ptr(self._d["node_center"]), _abi.ptr(self._node_info), self.node_count,
                  levels.ctypes.data, len(levels) - 1, _abi.ptr(binom), bsz, _abi.ptr(mom),
                  _abi.ptr(prime), _abi.stream())
        self._cart = (mom, prime)
        return self._cart

    def _cart_scaled(self) -> torch.Tensor:
        """FP32 m'_a / h^|a| of every cell, the mixed-precision Cartesian M2L's moments
        (kept with the moments until `set_charges`)."""
        cached = self.__dict__.get("_cart_f32")
        if cached is not None:
            return cached
        _, prime = self._cart_moments()
        out = torch.empty_like(prime, dtype=torch.float32)
        _abi.call("td3_cart_scale_prime", self.order, _abi.ptr(prime),
                  _abi.ptr(self._d["node_half"]), self.node_count, _abi.ptr(out), _abi.stream())
        self._cart_f32 = out
        return out

    def _cart_locals(self, target, m2l_rows, softening, precision=_abi.FP64):
        """Cartesian local expansions of every cell of ``target``: M2L (FP64, or FP32 on
        radius-scaled terms in mixed precision up to order 12) then L2L."""
        mrow, mcol, mdst = m2l_rows
        nc = (self.order + 1) * (self.order + 2) // 2
        dev = self._d["pos"].device
        local = torch.zeros((target.node_count, nc), dtype=torch.float64, device=dev)
        _, prime = self._cart_moments()
        nbytes = _abi.lib().td3_cart_m2l_workspace_bytes(int(mcol.shape[0]), self.order,
                                                         target.node_count)
        ws = _abi.workspace(nbytes)
        if precision != _abi.FP64 and self.order <= 12:
            _abi.call("td3_cart_m2l_f32", self.order, _abi.ptr(mcol), _abi.ptr(mdst),
                      _abi.ptr(mrow), 0, target.node_count, _abi.ptr(target._d["node_center"]),
                      _abi.ptr(target._d["node_half"]), _abi.ptr(self._d["node_center"]),
                      _abi.ptr(self._d["node_half"]), _abi.ptr(self._cart_scaled()),
                      float(softening), _abi.ptr(local), _abi.ptr(ws), nbytes, _abi.stream())
        else:
            _abi.call("td3_cart_m2l", self.order, _abi.ptr(mcol), _abi.ptr(mdst),
                      _abi.ptr(mrow), 0, target.node_count, _abi.ptr(target._d["node_center"]),
                      _abi.ptr(self._d["node_center"]), _abi.ptr(prime), float(softening),
                      _abi.ptr(local), _abi.ptr(ws), nbytes, _abi.stream())
        bsz = 2 * self.order + 1
        binom = _device_table("binomial", _binomial_table, bsz)
        levels = np.ascontiguousarray(target.levels, dtype=np.int64)
        _abi.call("td3_cart_l2l", self.order, _abi.ptr(target._d["node_center"]),
                  _abi.ptr(target._node_info), levels.ctypes.data, len(levels) - 1,
                  _abi.ptr(binom), bsz, _abi.ptr(local), _abi.stream())
        return local

    def cartesian_locals(self, theta: float = 0.5, softening: float = 0.0):
        """The Coulomb FMM's Cartesian local expansions (C, nc) after L2L."""
        local = self._cart_locals(self, self._lists(float(theta))[0], softening)
        return _out(local, self._like_torch)

    def _coulomb_fmm(self, theta, softening, potential, precision):
        """flat_tree.py:1360-1428: the Coulomb FMM with Cartesian expansions of the softened
        kernel (softened in the far field too)."""
        prec = _precision(precision or self.precision)
        m2l_rows, (prow, pcol, _), counts = self._lists(theta)
        local = self._cart_locals(self, m2l_rows, softening, prec)
        n = self._d["pos"].shape[1]
        dev = self._d["pos"].device
        phi = torch.zeros(n, dtype=torch.float64, device=dev)
        field = torch.zeros((2, n), dtype=torch.float64, device=dev)
        status = torch.zeros(1, dtype=torch.int64, device=dev)
        _abi.call(
            "td3_fmm_eval", _abi.COULOMB, prec, self.order, _abi.ptr(self._sorted_pos),
            _abi.ptr(self._scharge), _abi.ptr(self._d["perm"]), n,
            _abi.ptr(self._d["node_center"]), _abi.ptr(self._d["node_half"]),
            _abi.ptr(self._node_info), self.node_count, _abi.ptr(local), _abi.ptr(local),
            _abi.ptr(prow), _abi.ptr(pcol), _abi.ptr(self._leaf_index()),
            _abi.ptr(self._sources4()), float(softening), 1 if potential else 0,
            _abi.ptr(phi), _abi.ptr(field), _abi.ptr(status), _abi.stream(),
        )
        _check(status)
        self.fmm_pairs = counts
        return _out(phi, self._like_torch), _out(field, self._like_torch)

    # ---------------------------------------------------------------- fmm_at
    def _target_plan(self, targets: torch.Tensor, theta: float, leaf: int):
        """flat_tree.py:1124-1149: the target tree and the dual lists, kept while the
        targets are the same."""
        key = ("targets", theta, leaf)
        cached = self._traversals.get(key)
        if (cached is not None and cached[0].shape == targets.shape
                and bool(torch.equal(cached[0], targets))):
            return cached[1], cached[2]
        from .quad import Quad

        m = targets.shape[1]
        lo, hi = (v.cpu().numpy() for v in torch.aminmax(targets, dim=1))
        centre = 0.5 * (lo + hi)
        half = 0.5 * float(np.max(hi - lo)) * (1 + 1e-12) + 1e-300
        tree = FlatQuadtree(targets, torch.zeros(m, dtype=torch.float64, device=targets.device),
                            Quad(float(centre[0]), float(centre[1]), half), leaf_capacity=leaf,
                            max_depth=self.max_depth, kernel="coulomb")
        dev = targets.device
        ct, cs = tree.node_count, self.node_count
        caps = [16 * (ct + cs) + 64, 8 * (ct + cs) + 64, 10 * (ct + cs) + 64]
        out = np.zeros(4, dtype=np.int64)
        while True:
            m2l = torch.empty((caps[0], 2), dtype=torch.int32, device=dev)
            p2p = torch.empty((caps[1], 2), dtype=torch.int32, device=dev)
            nbytes = _abi.lib().td3_traverse_workspace_bytes(caps[2])
            ws = _abi.workspace(nbytes)
            code = _abi.lib().td3_dual_traverse(
                _abi.ptr(tree._d["node_center"]), _abi.ptr(tree._d["node_half"]),
                _abi.ptr(tree._node_info), ct, tree._depth, _abi.ptr(self._d["node_center"]),
                _abi.ptr(self._d["node_half"]), _abi.ptr(self._node_info), cs, self._depth,
                float(theta), caps[0], caps[1], caps[2], _abi.ptr(m2l), _abi.ptr(p2p),
                _abi.ptr(ws), nbytes, out.ctypes.data, _abi.stream(),
            )
            if code == _abi.TD3_ERR_CAPACITY:
                caps = [max(2 * caps[i], int(out[i]) + 64) for i in range(3)]
                continue
            _abi.check(code, "td3_dual_traverse")
            break
        rows = []
        for pairs, count in ((m2l, int(out[0])), (p2p, int(out[1]))):
            row = torch.empty(ct + 1, dtype=torch.int32, device=dev)
            col = torch.empty(max(count, 1), dtype=torch.int32, device=dev)
            dst = torch.empty(max(count, 1), dtype=torch.int32, device=dev)
            nbytes = _abi.lib().td3_lists_workspace_bytes(ct, count, 2)
            ws = _abi.workspace(nbytes)
            _abi.call("td3_lists", _abi.ptr(pairs), count, 2, ct, cs, _abi.ptr(row),
                      _abi.ptr(col), _abi.ptr(dst), _abi.ptr(ws), nbytes, _abi.stream())
            rows.append((row, col, dst))
        lists = (rows[0], rows[1], (int(out[0]), int(out[1])))
        self._traversals[key] = (targets.clone(), tree, lists)
        self.plan_builds += 1
        return tree, lists

    def fmm_at(self, targets, *, exclude=None, theta: float = 0.5, softening: float = 0.0,
               potential: bool = True, leaf_capacity=None, precision=None,
               jacobian: bool = False):
        """Potential ``(M,)`` and field ``(2, M)`` at ``targets`` ``(2, M)`` by the FMM.

        flat_tree.py:1152-1316: the targets get a tree of their own (``leaf_capacity``
        per leaf, by default the sources'), a dual traversal pairs its cells with the
        source cells, and local expansions run down the target tree to the targets.
        ``exclude[i]`` is a source index skipped for target ``i`` in the exact sums.
        ``jacobian`` and dipoles as in `fmm` (``(3, M)`` at the targets).
        """
        from ._utils import _finite

        theta = _finite("theta", theta, minimum=0)
        if not 0 < theta < 1:
            raise ValueError("FMM theta must lie in (0, 1)")
        softening = _finite("softening", softening, minimum=0)
        like_torch = _is_torch(targets) or self._like_torch
        t = _points(targets, "targets")
        if not bool(torch.isfinite(t).all()):
            raise ValueError("targets must be finite")
        m, n = t.shape[1], self._d["pos"].shape[1]
        _check_jacobian(self, jacobian, self.__dict__.get("_sdip"))
        phi = torch.zeros(m, dtype=torch.float64, device=t.device)
        field = torch.zeros((2, m), dtype=torch.float64, device=t.device)
        jac = torch.zeros((3, m), dtype=torch.float64, device=t.device) if jacobian else None
        if not m or not n:
            return self._fmm_out(phi, field, jac, like_torch)
        if exclude is None:
            ex = torch.full((m,), -1, dtype=torch.int64, device=t.device)
        else:
            ex = _device(exclude, torch.int64).reshape(-1)
        leaf = self.leaf_capacity if leaf_capacity is None else int(leaf_capacity)
        tree, (m2l_rows, (prow, pcol, _), counts) = self._target_plan(t, theta, leaf)
        self.fmm_at_pairs = counts
        if self.kernel != "log":
            return self._coulomb_fmm_at(t, ex, tree, m2l_rows, prow, pcol, softening, potential,
                                        precision, like_torch)
        prec = _precision(precision or self.precision)
        lre, lim = self._local_expansions(theta, prec, target=tree, m2l_rows=m2l_rows)
        status = torch.zeros(1, dtype=torch.int64, device=t.device)
        td = tree._d
        sdip = self.__dict__.get("_sdip")
        if jacobian or sdip is not None:
            _abi.call(
                "td3_log_fmm_at_eval_ext", self.order, _abi.ptr(tree._sorted_pos),
                _abi.ptr(td["perm"]), _abi.ptr(ex), m, _abi.ptr(td["node_center"]),
                _abi.ptr(td["node_half"]), _abi.ptr(tree._node_info), tree.node_count,
                _abi.ptr(lre),
                _abi.ptr(lim), _abi.ptr(prow), _abi.ptr(pcol), _abi.ptr(self._d["perm"]), n,
                _abi.ptr(self._node_info), _abi.ptr(self._sources4()), _abi.ptr(sdip),
                softening, _abi.ptr(phi), _abi.ptr(field), _abi.ptr(jac), _abi.ptr(status),
                _abi.stream(),
            )
            _check(status)
            if not potential:
                phi.zero_()
            self.fmm_at_pairs = counts
            return self._fmm_out(phi, field, jac, like_torch)
        _abi.call(
            "td3_fmm_at_eval", _abi.LOG, prec, self.order, _abi.ptr(tree._sorted_pos),
            _abi.ptr(td["perm"]),
            _abi.ptr(ex), m, _abi.ptr(td["node_center"]), _abi.ptr(td["node_half"]),
            _abi.ptr(tree._node_info), tree.node_count, _abi.ptr(tree._leaf_index()),
            _abi.ptr(lre), _abi.ptr(lim),
            _abi.ptr(prow), _abi.ptr(pcol), _abi.ptr(self._sorted_pos), _abi.ptr(self._scharge),
            _abi.ptr(self._d["perm"]), n, _abi.ptr(self._node_info),
            _abi.ptr(self._sources4()), softening,
            1 if potential else 0, _abi.ptr(phi), _abi.ptr(field), _abi.ptr(status),
            _abi.stream(),
        )
        _check(status)
        self.fmm_at_pairs = counts
        return _out(phi, like_torch), _out(field, like_torch)

    def _coulomb_fmm_at(self, t, ex, tree, m2l_rows, prow, pcol, softening, potential,
                        precision, like_torch):
        """fmm_at's Coulomb branch (flat_tree.py:1247-1295)."""
        prec = _precision(precision or self.precision)
        local = self._cart_locals(tree, m2l_rows, softening, prec)
        m, n = t.shape[1], self._d["pos"].shape[1]
        phi = torch.zeros(m, dtype=torch.float64, device=t.device)
        field = torch.zeros((2, m), dtype=torch.float64, device=t.device)
        status = torch.zeros(1, dtype=torch.int64, device=t.device)
        td = tree._d
        _abi.call(
            "td3_fmm_at_eval", _abi.COULOMB, prec, self.order, _abi.ptr(tree._sorted_pos),
            _abi.ptr(td["perm"]), _abi.ptr(ex), m, _abi.ptr(td["node_center"]),
            _abi.ptr(td["node_half"]), _abi.ptr(tree._node_info), tree.node_count,
            _abi.ptr(tree._leaf_index()), _abi.ptr(local), _abi.ptr(local), _abi.ptr(prow),
            _abi.ptr(pcol),
            _abi.ptr(self._sorted_pos), _abi.ptr(self._scharge), _abi.ptr(self._d["perm"]), n,
            _abi.ptr(self._node_info), _abi.ptr(self._sources4()), float(softening),
            1 if potential else 0, _abi.ptr(phi),
            _abi.ptr(field), _abi.ptr(status), _abi.stream(),
        )
        _check(status)
        return _out(phi, like_torch), _out(field, like_torch)


def _check_jacobian(tree, jacobian, sdip) -> None:
    if jacobian and tree.kernel != "log":
        raise ValueError("jacobian needs the log kernel")
    if jacobian and sdip is not None:
        raise ValueError("jacobian of the field of dipoles is not provided")


def _m2l_coefficients(order: int, symmetric: bool) -> np.ndarray:
    """coeff[k, l] of the M2L (flat_tree.py:234-245): coeff[0, l] = 1/l, coeff[k, 0] = 1/k,
    coeff[k, l] = C(k + l - 1, l) / k, zero beyond the truncation (k <= p, or k <= p - l
    when symmetric)."""
    binomial = _binomial_table(2 * order + 1)
    coeff = np.zeros((order + 1, order + 1))
    for m in range(1, order + 1):
        coeff[0, m] = 1.0 / m
    for k in range(1, order + 1):
        coeff[k, 0] = 1.0 / k
        for m in range(1, order + 1):
            if k <= (order - m if symmetric else order):
                coeff[k, m] = binomial[k + m - 1, m] / k
    return coeff
