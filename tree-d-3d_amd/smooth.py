"""Smooth-membership line-charge FMM on the GPU: forces that are the exact gradient of an
energy, with exactly antisymmetric pair forces (the reference's smooth.py:1-476).

Same interface as the reference: ``BoxTree`` (static square cells, kept across time
steps), ``SmoothFMM`` (``energy``, ``fields()``, ``potential_at``), ``direct_energy``,
``direct_fields`` and ``comoving_fields``; units of ``flat_tree`` (potential of a line
charge ``-2 q ln r``, field ``2 q d / r^2``, no ``k_e``).  A particle belongs to every
leaf whose box grown by ``band`` holds it, with a weight that is a product of quintic edge
steps; ``band = 0`` is the ordinary tree.  Softening is the compact cloud of radius
``softening``: beyond it the kernel is the bare one, so the far field is exact.

Everything runs in FP64 through libtd3_hip.so: td3_box_tree_build (the cells),
td3_box_leaf_order, td3_smooth_members (memberships and their layout by leaf),
td3_moments / td3_rows / td3_scale_moments / td3_m2l / td3_l2l of the log FMM on that
layout, td3_smooth_eval, td3_smooth_combine, and td3_compact_direct for the direct sums.
Results come back in the caller's array type (NumPy for NumPy inputs, device tensors for
torch inputs).  Membership arrays (``part``, ``leaf``, ``w``, ...) hold the reference's
set of memberships ordered by leaf (the reference orders them by descent level).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _abi
from ._utils import _device, _finite, _is_torch, _out, _points
from .flat_tree import MAX_ORDER, _binomial_table, _device_table, _m2l_coefficients

SLOTS = 4
BOX_MAX_DEPTH = 40  # smooth.py:97

#: first capacities of the interaction rows per cell and the slot width of td3_rows
_ROW_CAPS = (96, 40)
_ROW_WIDTH = 128


def _status_error(code: int, what: str) -> None:
    if code > 0:
        raise ValueError(f"{what}: particles {code - 1} and another coincide (softening 0)")


class BoxTree:
    """Square cells of a quadtree over ``root``, split to hold ``leaf_capacity`` points
    (smooth.py:83-121; ``center`` complex, ``half``, ``parent``, ``child``, ``depth``,
    ``leaves``, ``min_leaf_half``, ``root``)."""

    def __init__(self, pos, root, *, leaf_capacity: int = 16) -> None:
        for name, value in (("xmid", root.xmid), ("ymid", root.ymid)):
            _finite(name, value)
        _finite("half", root.half, positive=True)
        if isinstance(leaf_capacity, bool) or int(leaf_capacity) != leaf_capacity \
                or leaf_capacity < 1:
            raise ValueError("leaf_capacity must be a positive integer")
        self.root = root
        self.leaf_capacity = int(leaf_capacity)
        self._like_torch = _is_torch(pos)
        pts = _points(pos, "positions")
        if pts.shape[1] and not bool(torch.isfinite(pts).all()):
            raise ValueError("Particle position must be finite")
        self._build(pts)
        self._order()
        self._host: dict = {}

    def _build(self, pts: torch.Tensor) -> None:
        dev = pts.device
        n = pts.shape[1]
        xmid, ymid, half = float(self.root.xmid), float(self.root.ymid), float(self.root.half)
        f64 = dict(dtype=torch.float64, device=dev)
        i64 = dict(dtype=torch.int64, device=dev)
        if n == 0:
            self.node_count = 1
            self.levels = np.array([0, 1], dtype=np.int64)
            self._center = torch.tensor([[xmid, ymid]], **f64)
            self._half = torch.tensor([half], **f64)
            self._child = torch.full((1,), -1, **i64)
            self._depth = torch.zeros(1, **i64)
            self._info = torch.tensor([[0, 0, -1, 0]], dtype=torch.int32, device=dev)
            return
        capacity = 4 * (2 * n // self.leaf_capacity + 16)
        out = np.zeros(3, dtype=np.int64)
        levels_host = np.zeros(67, dtype=np.int64)
        while True:
            center = torch.empty((capacity, 2), **f64)
            half_arr = torch.empty(capacity, **f64)
            start, end = torch.empty(capacity, **i64), torch.empty(capacity, **i64)
            child, depth = torch.empty(capacity, **i64), torch.empty(capacity, **i64)
            levels = torch.zeros(67, **i64)
            perm = torch.empty(n, **i64)
            sorted_pos = torch.empty((2, n), **f64)
            info = torch.empty((capacity, 4), dtype=torch.int32, device=dev)
            nbytes = _abi.lib().td3_tree_workspace_bytes(n, capacity)
            ws = _abi.workspace(nbytes)
            code = _abi.lib().td3_box_tree_build(
                _abi.ptr(pts), n, xmid, ymid, half, self.leaf_capacity, BOX_MAX_DEPTH, capacity,
                _abi.ptr(center), _abi.ptr(half_arr), _abi.ptr(start), _abi.ptr(end),
                _abi.ptr(child), _abi.ptr(depth), _abi.ptr(levels), _abi.ptr(perm),
                _abi.ptr(sorted_pos), _abi.ptr(info), _abi.ptr(ws), nbytes, out.ctypes.data,
                levels_host.ctypes.data, _abi.stream(),
            )
            if code == _abi.TD3_ERR_CAPACITY:
                capacity = max(2 * capacity, int(out[2]))
                continue
            _abi.check(code, "td3_box_tree_build")
            break
        count, n_levels = int(out[0]), int(out[1])
        self.node_count = count
        self.levels = levels_host[: n_levels + 1].copy()
        self._center = center[:count]
        self._half = half_arr[:count]
        self._child = child[:count]
        self._depth = depth[:count]
        self._info = info[:count]

    def _order(self) -> None:
        """Depth-first leaf ranks (td3_box_leaf_order) and the smallest leaf half-width."""
        c = self.node_count
        dev = self._center.device
        self._rank = torch.empty(c, dtype=torch.int32, device=dev)
        self._span = torch.empty(c, dtype=torch.int32, device=dev)
        levels = np.ascontiguousarray(self.levels, dtype=np.int64)
        _abi.call("td3_box_leaf_order", _abi.ptr(self._info), c, levels.ctypes.data,
                  len(levels) - 1, _abi.ptr(self._rank), _abi.ptr(self._span), _abi.stream())
        leaf = self._child < 0
        both = torch.stack((self._span[0].to(torch.float64),
                            torch.where(leaf, self._half, torch.inf).min())).cpu()
        self.leaf_count = int(both[0])
        self.min_leaf_half = float(both[1])

    # ---------------------------------------------------------------- reference arrays
    def _array(self, name: str, make):
        if self._like_torch:
            return make()
        if name not in self._host:
            self._host[name] = make().cpu().numpy()
        return self._host[name]

    @property
    def center(self):
        return self._array("center", lambda: torch.complex(self._center[:, 0].contiguous(),
                                                           self._center[:, 1].contiguous()))

    @property
    def half(self):
        return self._array("half", lambda: self._half)

    @property
    def child(self):
        return self._array("child", lambda: self._child)

    @property
    def depth(self):
        return self._array("depth", lambda: self._depth)

    @property
    def parent(self):
        def make():
            parent = torch.full_like(self._child, -1)
            inner = torch.nonzero(self._child >= 0).reshape(-1)
            for k in range(4):
                parent[self._child[inner] + k] = inner
            return parent
        return self._array("parent", make)

    @property
    def leaves(self):
        return self._array("leaves", lambda: torch.nonzero(self._child < 0).reshape(-1))

    def _check_band(self, band: float) -> float:
        band = _finite("band", band)
        if band > 0.0 and band > self.min_leaf_half:
            raise ValueError("band must not exceed the smallest leaf half-width")
        return band

    def _memberships(self, pts: torch.Tensor, charge: torch.Tensor, band: float) -> dict:
        """td3_smooth_members: slots, the memberships sorted by leaf, and their node_info."""
        n = pts.shape[1]
        dev = pts.device
        cap = SLOTS * n
        f64 = dict(dtype=torch.float64, device=dev)
        i32 = dict(dtype=torch.int32, device=dev)
        d = dict(
            slot_leaf=torch.empty((n, SLOTS), **i32),
            slot_weight=torch.empty((3, n, SLOTS), **f64),
            slot_member=torch.empty((n, SLOTS), **i32),
            member_pos=torch.empty((2, cap), **f64),
            member_charge=torch.empty(cap, **f64),
            member_src=torch.empty((cap, 4), **f64),
            member_leaf=torch.empty(cap, **i32),
            member_info=torch.empty((self.node_count, 4), **i32),
            member_count=torch.zeros(1, dtype=torch.int64, device=dev),
            status=torch.zeros(2, dtype=torch.int64, device=dev),
        )
        nbytes = _abi.lib().td3_smooth_members_workspace_bytes(n, self.leaf_count)
        ws = _abi.workspace(nbytes)
        xmid, ymid, half = float(self.root.xmid), float(self.root.ymid), float(self.root.half)
        _abi.call(
            "td3_smooth_members", _abi.ptr(pts), _abi.ptr(charge), n, _abi.ptr(self._center),
            _abi.ptr(self._half), _abi.ptr(self._info), self.node_count, _abi.ptr(self._rank),
            _abi.ptr(self._span), self.leaf_count, xmid, ymid, half, float(band),
            _abi.ptr(d["slot_leaf"]), _abi.ptr(d["slot_weight"]), _abi.ptr(d["slot_member"]),
            _abi.ptr(d["member_pos"]), _abi.ptr(d["member_charge"]), _abi.ptr(d["member_src"]),
            _abi.ptr(d["member_leaf"]), _abi.ptr(d["member_info"]), _abi.ptr(d["member_count"]),
            _abi.ptr(d["status"]), _abi.ptr(ws), nbytes, _abi.stream(),
        )
        return d

    def memberships(self, z, band: float):
        """(particle, leaf, weight, dw/dx, dw/dy) for every particle-leaf pair with w > 0
        (smooth.py:122-193), ordered by particle.  ``z`` complex (N,)."""
        like_torch = _is_torch(z)
        zt = z if like_torch else torch.from_numpy(np.ascontiguousarray(np.asarray(z, dtype=complex)))
        zt = zt.to(device=_abi.device(), dtype=torch.complex128).reshape(-1)
        band = self._check_band(band)
        n = zt.shape[0]
        if n == 0:
            empty_i = torch.zeros(0, dtype=torch.int64, device=zt.device)
            empty_f = torch.zeros(0, dtype=torch.float64, device=zt.device)
            return tuple(_out(t, like_torch) for t in (empty_i, empty_i, empty_f, empty_f, empty_f))
        pts = torch.stack((zt.real, zt.imag)).contiguous()
        d = self._memberships(pts, torch.zeros(n, dtype=torch.float64, device=pts.device), band)
        _status_members(d["status"])
        leaf = d["slot_leaf"].reshape(-1)
        keep = leaf >= 0
        part = torch.arange(n, device=pts.device).repeat_interleave(SLOTS)[keep]
        sw = d["slot_weight"].reshape(3, -1)[:, keep]
        out = (part, leaf[keep].to(torch.int64), sw[0], sw[1], sw[2])
        return tuple(_out(t.contiguous(), like_torch) for t in out)


def _status_members(status: torch.Tensor) -> None:
    code = int(status[0].item())
    if code > 0:
        raise RuntimeError(f"particle {code - 1} lies in more than four grown cells of one level")


def _rows(tree: BoxTree, info: torch.Tensor, theta: float):
    """td3_rows on the memberships' node_info (empty cells are skipped: the reference's
    `occupied`), regrown until they fit."""
    dev = info.device
    c = tree.node_count
    levels = np.ascontiguousarray(tree.levels, dtype=np.int64)
    caps = [f * c + 64 for f in _ROW_CAPS] + [_ROW_WIDTH]
    out = np.zeros(4, dtype=np.int64)
    i32 = dict(dtype=torch.int32, device=dev)
    while True:
        mrow, prow = torch.empty(c + 1, **i32), torch.empty(c + 1, **i32)
        mcol, mdst = torch.empty(caps[0], **i32), torch.empty(caps[0], **i32)
        pcol = torch.empty(caps[1], **i32)
        nbytes = _abi.lib().td3_rows_workspace_bytes(levels.ctypes.data, len(levels) - 1, caps[2])
        ws = _abi.workspace(nbytes)
        code = _abi.lib().td3_rows(
            _abi.ptr(tree._center), _abi.ptr(tree._half), _abi.ptr(info), levels.ctypes.data,
            len(levels) - 1, float(theta), caps[0], caps[1], caps[2], _abi.ptr(mrow),
            _abi.ptr(mcol), _abi.ptr(mdst), _abi.ptr(prow), _abi.ptr(pcol), _abi.ptr(ws), nbytes,
            out.ctypes.data, None, _abi.stream(),
        )
        if code == _abi.TD3_ERR_CAPACITY:
            caps = [max(2 * caps[0], int(out[0]) + 64), max(2 * caps[1], int(out[1]) + 64),
                    max(caps[2], int(out[3]) + 8) if out[3] > 1 else caps[2]]
            continue
        _abi.check(code, "td3_rows")
        break
    n_m2l, n_p2p = int(out[0]), int(out[1])
    return (mrow, mcol[: max(n_m2l, 1)], mdst[: max(n_m2l, 1)], n_m2l), \
        (prow, pcol[: max(n_p2p, 1)], n_p2p)


class SmoothFMM:
    """Potential energy and exact-gradient forces of line charges on a `BoxTree`
    (smooth.py:242-428; same constructor and defaults)."""

    def __init__(self, tree: BoxTree, pos, charge, *, order: int = 12, theta: float = 0.5,
                 band: float = 0.0, symmetric: bool = True, softening: float = 0.0) -> None:
        if isinstance(order, bool) or int(order) != order or not 1 <= order <= MAX_ORDER:
            raise ValueError(f"order must be an integer in [1, {MAX_ORDER}]")
        theta = _finite("theta", theta, positive=True)
        self.tree = tree
        self.order = int(order)
        self.theta = theta
        self.band = tree._check_band(band)
        self.symmetric = bool(symmetric)
        self.softening = _finite("softening", softening, minimum=0)
        self.eps2 = self.softening ** 2
        self._like_torch = _is_torch(pos)
        self._pos = _points(pos, "positions")
        self._q = _device(charge).reshape(-1)
        n = self._pos.shape[1]
        if self._q.shape[0] != n:
            raise ValueError("positions and charges must have the same length")
        if n and not bool(torch.isfinite(self._pos).all()):
            raise ValueError("Particle position must be finite")
        self._host: dict = {}
        self._compute()

    # ---------------------------------------------------------------- the computation
    def _compute(self) -> None:
        t, n = self.tree, self._pos.shape[1]
        dev = self._pos.device
        f64 = dict(dtype=torch.float64, device=dev)
        self.m2l_count = self.p2p_count = 0
        if n == 0:
            self._phi = torch.zeros(0, **f64)
            self._field = torch.zeros((2, 0), **f64)
            self._energy = torch.zeros(1, **f64)
            self._m = None
            return
        c, p1, cap = t.node_count, self.order + 1, SLOTS * n
        m = t._memberships(self._pos, self._q, self.band)
        self._m = m
        levels = np.ascontiguousarray(t.levels, dtype=np.int64)
        n_levels = len(levels) - 1
        bsz = 2 * self.order + 1
        binom = _device_table("binomial", _binomial_table, bsz)
        # moments of the weighted memberships below each cell (smooth.py:287-303)
        mre, mim = torch.empty((c, p1), **f64), torch.empty((c, p1), **f64)
        qabs = torch.empty(c, **f64)
        node_q, node_p, node_s = torch.empty(c, **f64), torch.empty((c, 2), **f64), \
            torch.empty((c, 3), **f64)
        _abi.call("td3_moments", _abi.LOG, self.order, _abi.ptr(m["member_pos"]),
                  _abi.ptr(m["member_charge"]), cap, _abi.ptr(t._center), _abi.ptr(m["member_info"]),
                  c, levels.ctypes.data, n_levels, _abi.ptr(binom), bsz, _abi.ptr(node_q),
                  _abi.ptr(qabs), _abi.ptr(node_p), _abi.ptr(node_s), _abi.ptr(mre), _abi.ptr(mim),
                  _abi.stream())
        self._mre, self._mim = mre, mim
        # the dual walk over occupied cells (smooth.py:246-277, :304)
        (mrow, mcol, mdst, n_m2l), (prow, pcol, n_p2p) = _rows(t, m["member_info"], self.theta)
        self._rows = (mrow, mcol, mdst, prow, pcol)
        self.m2l_count, self.p2p_count = n_m2l, n_p2p
        # M2L both ways (smooth.py:306-331) on radius-scaled moments, then L2L (:332-342)
        are, aim = torch.empty((c, p1), **f64), torch.empty((c, p1), **f64)
        _abi.call("td3_scale_moments", _abi.FP64, _abi.ptr(mre), _abi.ptr(mim), _abi.ptr(t._half),
                  c, self.order, _abi.ptr(are), _abi.ptr(aim), _abi.stream())
        lre, lim = torch.zeros((c, p1), **f64), torch.zeros((c, p1), **f64)
        coeff = _device_table("m2l", _m2l_coefficients, self.order, self.symmetric)
        if n_m2l:
            if p1 <= 16:
                nbytes, ws = 0, None  # the matrix-core path, a wavefront per row
            else:
                nbytes = _abi.lib().td3_m2l_workspace_bytes(int(mcol.shape[0]), self.order,
                                                            _abi.FP64)
                ws = _abi.workspace(nbytes)
            _abi.call("td3_m2l", _abi.FP64, self.order, _abi.ptr(mcol), _abi.ptr(mdst),
                      _abi.ptr(mrow), 0, c, _abi.ptr(t._center), _abi.ptr(t._half),
                      _abi.ptr(t._center), _abi.ptr(t._half), _abi.ptr(are), _abi.ptr(aim),
                      _abi.ptr(coeff), _abi.ptr(lre), _abi.ptr(lim), _abi.ptr(ws), nbytes,
                      _abi.stream())
        _abi.call("td3_l2l", self.order, _abi.ptr(t._center), _abi.ptr(t._half),
                  _abi.ptr(m["member_info"]), levels.ctypes.data, n_levels, _abi.ptr(binom), bsz,
                  _abi.ptr(lre), _abi.ptr(lim), _abi.stream())
        self._lre, self._lim = lre, lim
        # Phi_m, G_m per membership (smooth.py:344-380), then per particle (:410-422)
        phi_m, g_m = torch.empty(cap, **f64), torch.empty((2, cap), **f64)
        _abi.call("td3_smooth_eval", self.order, _abi.ptr(m["member_src"]),
                  _abi.ptr(m["member_leaf"]), cap, _abi.ptr(m["member_count"]), _abi.ptr(t._center),
                  _abi.ptr(t._half), _abi.ptr(m["member_info"]), _abi.ptr(lre), _abi.ptr(lim),
                  _abi.ptr(prow), _abi.ptr(pcol), self.softening, _abi.ptr(phi_m), _abi.ptr(g_m),
                  _abi.ptr(m["status"][1:]), _abi.stream())
        self._phi_m, self._g_m = phi_m, g_m
        self._phi, self._field = torch.empty(n, **f64), torch.empty((2, n), **f64)
        self._energy = torch.empty(1, **f64)
        nbytes = _abi.lib().td3_smooth_combine_workspace_bytes(n)
        ws = _abi.workspace(nbytes)
        _abi.call("td3_smooth_combine", n, cap, _abi.ptr(self._q), _abi.ptr(m["slot_member"]),
                  _abi.ptr(m["slot_weight"]), _abi.ptr(phi_m), _abi.ptr(g_m), _abi.ptr(self._phi),
                  _abi.ptr(self._field), _abi.ptr(self._energy), _abi.ptr(ws), nbytes,
                  _abi.stream())
        host = torch.cat((m["status"], m["member_count"])).cpu().numpy()
        _status_members(host[:1])
        _status_error(int(host[1]), "SmoothFMM")
        self.membership_count = int(host[2])

    # ---------------------------------------------------------------- results
    @property
    def energy(self) -> float:
        """U = -2 sum over unordered routes (smooth.py:410-413)."""
        return float(self._energy.item())

    def fields(self):
        """Potential ``(N,)`` and field ``(2, N)`` with ``F_i = q_i E_i = -dU/dx_i``
        (smooth.py:415-422)."""
        return _out(self._phi, self._like_torch), _out(self._field, self._like_torch)

    def potential_at(self, points):
        """Direct-sum potential at external points (smooth.py:424-428)."""
        pts = _points(points, "points")
        phi = _compact_direct(pts, self._pos, self._q, self.softening, potential=True,
                              field=False)[0]
        return _out(phi, _is_torch(points))

    # ---------------------------------------------------------------- membership arrays
    def _array(self, name: str, make):
        if name not in self._host:
            self._host[name] = make()
        t = self._host[name]
        return t if self._like_torch else t.cpu().numpy()

    def _sorted(self, source: torch.Tensor) -> torch.Tensor:
        """Per-slot values (N, S) moved to the sorted membership order."""
        m = self._m
        j = m["slot_member"].reshape(-1)
        keep = j >= 0
        out = torch.empty(self.membership_count, dtype=source.dtype, device=source.device)
        out[j[keep].to(torch.int64)] = source.reshape(-1)[keep]
        return out

    @property
    def part(self):
        def make():
            if self._m is None:
                return torch.zeros(0, dtype=torch.int64, device=self._pos.device)
            return self._m["member_src"][: self.membership_count, 3].to(torch.int64)
        return self._array("part", make)

    @property
    def leaf(self):
        def make():
            if self._m is None:
                return torch.zeros(0, dtype=torch.int64, device=self._pos.device)
            return self._m["member_leaf"][: self.membership_count].to(torch.int64)
        return self._array("leaf", make)

    @property
    def w(self):
        def make():
            if self._m is None:
                return torch.zeros(0, dtype=torch.float64, device=self._pos.device)
            return self._sorted(self._m["slot_weight"][0])
        return self._array("w", make)

    @property
    def dw(self):
        def make():
            if self._m is None:
                return torch.zeros(0, dtype=torch.complex128, device=self._pos.device)
            sw = self._m["slot_weight"]
            return torch.complex(self._sorted(sw[1]), self._sorted(sw[2]))
        return self._array("dw", make)

    @property
    def phi_m(self):
        return self._array("phi_m", lambda: self._phi_m[: self.membership_count].clone()
                           if self._m is not None else torch.zeros(0, dtype=torch.float64))

    @property
    def g_m(self):
        def make():
            if self._m is None:
                return torch.zeros(0, dtype=torch.complex128)
            k = self.membership_count
            return torch.complex(self._g_m[0, :k].contiguous(), self._g_m[1, :k].contiguous())
        return self._array("g_m", make)

    @property
    def moments(self):
        """Unscaled multipole moments a_k = sum q w (z - c)^k of every cell, (C, order+1)."""
        return self._array("moments", lambda: torch.complex(self._mre, self._mim))

    @property
    def local(self):
        """Unscaled local expansions b_l of every cell after L2L, (C, order+1)."""
        def make():
            lre, lim = self._lre.clone(), self._lim.clone()
            _abi.call("td3_unscale_locals", _abi.ptr(self.tree._half), self.tree.node_count,
                      self.order, _abi.ptr(lre), _abi.ptr(lim), _abi.stream())
            return torch.complex(lre, lim)
        return self._array("local", make)

    @property
    def m2l_pairs(self):
        """The well-separated cell pairs (K, 2), each once (smooth.py:305)."""
        def make():
            if self._m is None or not self.m2l_count:
                return torch.zeros((0, 2), dtype=torch.int64, device=self._pos.device)
            _, mcol, mdst, _, _ = self._rows
            col, dst = mcol.to(torch.int64), mdst.to(torch.int64)
            once = dst < col
            return torch.stack((dst[once], col[once]), dim=1)
        return self._array("m2l_pairs", make)

    @property
    def p2p_pairs(self):
        """The near leaf pairs (K, 2), each once, a leaf with itself included (smooth.py:305)."""
        def make():
            if self._m is None or not self.p2p_count:
                return torch.zeros((0, 2), dtype=torch.int64, device=self._pos.device)
            _, _, _, prow, pcol = self._rows
            counts = (prow[1:] - prow[:-1]).to(torch.int64)
            row = torch.repeat_interleave(torch.arange(len(counts), device=prow.device), counts)
            col = pcol[: row.shape[0]].to(torch.int64)
            once = row <= col
            return torch.stack((row[once], col[once]), dim=1)
        return self._array("p2p_pairs", make)


def _compact_direct(targets: torch.Tensor, pos: torch.Tensor, charge: torch.Tensor,
                    softening: float, *, potential: bool, field: bool, self_pairs: bool = False,
                    energy: bool = False):
    m, n = targets.shape[1], pos.shape[1]
    dev = pos.device
    f64 = dict(dtype=torch.float64, device=dev)
    phi = torch.zeros(m, **f64) if potential else None
    fld = torch.zeros((2, m), **f64) if field else None
    u = torch.zeros(1, **f64) if energy else None
    if m == 0 or n == 0:
        return phi, fld, u
    status = torch.zeros(1, dtype=torch.int64, device=dev)
    nbytes = _abi.lib().td3_compact_direct_workspace_bytes(m)
    ws = _abi.workspace(nbytes)
    _abi.call("td3_compact_direct", _abi.ptr(targets), m, 1 if self_pairs else 0, _abi.ptr(pos),
              _abi.ptr(charge), n, float(softening), _abi.ptr(phi), _abi.ptr(fld), _abi.ptr(u),
              _abi.ptr(status), _abi.ptr(ws), nbytes, _abi.stream())
    _status_error(int(status.item()), "direct sum")
    return phi, fld, u


def _direct_inputs(pos, charge, softening):
    softening = _finite("softening", softening, minimum=0)
    pts = _points(pos, "positions")
    q = _device(charge).reshape(-1)
    if q.shape[0] != pts.shape[1]:
        raise ValueError("positions and charges must have the same length")
    return pts, q, softening


def direct_energy(pos, charge, softening: float = 0.0) -> float:
    """Exact pair energy, in the units of `SmoothFMM.energy` (smooth.py:196-201)."""
    pts, q, softening = _direct_inputs(pos, charge, softening)
    _, _, u = _compact_direct(pts, pts, q, softening, potential=False, field=False,
                              self_pairs=True, energy=True)
    return float(u.item())


def direct_fields(pos, charge, softening: float = 0.0):
    """Exact potential and field, in the units of `SmoothFMM.fields` (smooth.py:431-440)."""
    pts, q, softening = _direct_inputs(pos, charge, softening)
    phi, fld, _ = _compact_direct(pts, pts, q, softening, potential=True, field=True,
                                  self_pairs=True)
    like = _is_torch(pos)
    return _out(phi, like), _out(fld, like)


def comoving_fields(tree: BoxTree, pos, charge, mass, center, **options):
    """Energy, potential and field with the boxes attached to the centre of mass
    (smooth.py:453-476): the cells of ``tree`` are used in the frame that follows the
    centre of mass, ``U(x) = U_tree(x - (c(x) - center))``, and the field loses each
    particle's mass share of the net force (divided by its charge)."""
    like = _is_torch(pos)
    pts = _points(pos, "positions")
    q = _device(charge).reshape(-1)
    mass_t = _device(mass).reshape(-1)
    c0 = _device(center).reshape(-1)
    com = (mass_t * pts).sum(dim=1) / mass_t.sum()
    f = SmoothFMM(tree, (pts - (com - c0)[:, None]).contiguous(), q, **options)
    phi, field = f._phi, f._field
    net = (q * field).sum(dim=1)
    field = field - (mass_t / mass_t.sum())[None, :] * net[:, None] / q[None, :]
    return f.energy, _out(phi, like), _out(field.contiguous(), like)
