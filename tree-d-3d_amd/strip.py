"""Line charges between two grounded plates, by conformal mapping and the FMM, on the GPU.

The reference's ``GroundedStripField`` (strip.py:1-327): charges are infinite line charges
along z in the strip 0 < x < L between grounded plates at x = 0 and x = L.  The map
zeta = exp(i pi z / L) sends the strip onto the upper half-plane, where a grounded
boundary takes one image per charge (the conjugate point, opposite charge), so the field
at the charges is one free-space line-charge FMM over 2N sources in the mapped plane,
then mapped back: E_x - i E_y = 2 k_e W'(zeta) (i pi / L) zeta.  A charge's own term in
the mapped plane is not purely a self-interaction; its regular part (gradient
i pi / 2L, value ln|zeta'(z_i)|) is restored.  Softening is applied in the physical
plane to pairs closer than ``cutoff * softening`` (td3_soften_near: the softened minus
bare free-space kernel, owner-computes on the GPU).

Everything runs on the GPU: the map and its inverse on device tensors, the FMM (or the
walk for probes, or the direct sum for ``method="direct"``) and the close-pair
correction through libtd3_hip.so.
"""

from __future__ import annotations

import math

import numpy as np
import torch

from . import _abi
from ._utils import _device, _is_torch, _out, _points
from .constants import K_E
from .flat_tree import FlatQuadtree, direct_field
from .quad import Quad


def _root(xy: torch.Tensor, grow: float = 1e-9) -> Quad:
    """strip.py:153-158: the square around the points, grown by a relative 1e-9."""
    lo = xy.min(dim=1).values.cpu().numpy()
    hi = xy.max(dim=1).values.cpu().numpy()
    centre = (lo + hi) / 2
    half = float(np.max(hi - lo) / 2) * (1 + grow) + 1e-300
    return Quad(float(centre[0]), float(centre[1]), half)


class GroundedStripField:
    """Field and potential of line charges between plates at x = 0 and x = L
    (strip.py:100-327: same constructor, defaults, methods and errors)."""

    def __init__(self, width: float, *, order: int = 14, theta: float = 0.5,
                 leaf_capacity: int = 32, k_e: float = K_E, method: str = "fmm",
                 softening: float = 0.0, cutoff: float = 6.0) -> None:
        if not np.isfinite(width) or width <= 0:
            raise ValueError("width must be finite and positive")
        if method not in ("fmm", "direct"):
            raise ValueError("method must be 'fmm' or 'direct'")
        self.width = float(width)
        self.order = int(order)
        self.theta = float(theta)
        self.leaf_capacity = int(leaf_capacity)
        self.k_e = float(k_e)
        self.method = method
        if not np.isfinite(softening) or softening < 0:
            raise ValueError("softening must be finite and nonnegative")
        if not cutoff > 1:
            raise ValueError("cutoff must exceed 1")
        self.softening = float(softening)
        self.cutoff = float(cutoff)

    # ------------------------------------------------------------------ the map
    def _plane(self, positions: torch.Tensor):
        """(Re zeta, Im zeta) of zeta = exp(i pi z / L) for (2, N) points strictly inside."""
        if positions.shape[1] and bool(((positions[0] <= 0) | (positions[0] >= self.width)).any()):
            raise ValueError("Charges must lie strictly between the plates")
        k = math.pi / self.width
        mod = torch.exp(-k * positions[1])
        return mod * torch.cos(k * positions[0]), mod * torch.sin(k * positions[0])

    def to_plane(self, positions):
        """zeta = exp(i pi z / L) as a complex array (strip.py:136-141)."""
        p = _points(positions, "positions")
        re, im = self._plane(p)
        z = torch.complex(re, im)
        return z if _is_torch(positions) else z.cpu().numpy()

    def _sources(self, positions: torch.Tensor, charges: torch.Tensor):
        """strip.py:143-151: the charges and their images in the mapped plane."""
        re, im = self._plane(positions)
        q = charges.reshape(-1)
        if q.shape[0] != re.shape[0]:
            raise ValueError("positions and charges must have the same length")
        xy = torch.stack((torch.cat((re, re)), torch.cat((im, -im))))
        return (re, im), xy.contiguous(), torch.cat((q, -q)).contiguous()

    def _mapped_field(self, raw: torch.Tensor, re: torch.Tensor, im: torch.Tensor):
        """E_x - i E_y = raw (i pi / L) zeta, returned as (E_x + i E_y) parts (Ex, Ey)."""
        k = math.pi / self.width
        # w = raw_x - i raw_y; c = w * (i k) * zeta
        wr, wi = raw[0], -raw[1]
        ar, ai = -k * wi, k * wr  # w * (i k)
        cr = ar * re - ai * im
        ci = ar * im + ai * re
        return cr, ci

    # ------------------------------------------------------------------ softening
    def _soften(self, sources: torch.Tensor, charges: torch.Tensor, targets: torch.Tensor,
                exclude: torch.Tensor | None, phi: torch.Tensor, field: torch.Tensor,
                with_potential: bool, message: str) -> None:
        """Adds the softened minus bare free-space kernel of every source within
        cutoff * softening of a target (td3_soften_near), in the physical plane."""
        n, m = sources.shape[1], targets.shape[1]
        if not n or not m or self.softening <= 0:
            return
        tree = FlatQuadtree(sources, charges, _root(sources), leaf_capacity=8)
        ws_bytes = _abi.lib().td3_soften_near_workspace_bytes(tree.node_count)
        ws = _abi.workspace(ws_bytes)
        status = torch.zeros(1, dtype=torch.int64, device=sources.device)
        order = tree._d["perm"] if exclude is not None else None
        _abi.call("td3_soften_near", _abi.ptr(targets), m, _abi.ptr(order), _abi.ptr(exclude),
                  _abi.ptr(tree._sorted_pos), _abi.ptr(tree._scharge), _abi.ptr(tree._d["perm"]),
                  n, _abi.ptr(tree._d["node_center"]), _abi.ptr(tree._d["node_half"]),
                  _abi.ptr(tree._node_info), tree.node_count, self.cutoff * self.softening,
                  self.softening, 1 if with_potential else 0, _abi.ptr(phi), _abi.ptr(field),
                  _abi.ptr(status), _abi.ptr(ws), ws_bytes, _abi.stream())
        if int(status[0]):
            raise ValueError(message)

    # ------------------------------------------------------------------ results
    def field_at_charges(self, positions, charges, *, potential: bool = True):
        """Potential (V) and field (V/m, (2, N)) acting on every charge (strip.py:185-233):
        all other charges and every image, its own included (the plates' image force)."""
        like = _is_torch(positions) or _is_torch(charges)
        pos = _points(positions, "positions")
        q_in = _device(charges).reshape(-1)
        (re, im), xy, q = self._sources(pos, q_in)
        n = re.shape[0]
        if not n:
            return _out(torch.zeros(0, dtype=torch.float64, device=pos.device), like), \
                _out(torch.zeros((2, 0), dtype=torch.float64, device=pos.device), like)
        if self.method == "fmm":
            tree = FlatQuadtree(xy, q, _root(xy), kernel="log", order=self.order,
                                leaf_capacity=self.leaf_capacity)
            phi, raw = tree.fmm(theta=self.theta, potential=potential)
        else:
            own = torch.arange(2 * n, device=pos.device)
            phi, raw = direct_field(xy, xy, q, exclude=own, kernel="log")
        cr, ci = self._mapped_field(raw[:, :n], re, im)
        # the regular part of each charge's own mapped term: + 2 q (i pi / 2L)
        cr = cr.clone()
        ci = ci + 2 * q_in * (math.pi / (2 * self.width))
        field = torch.stack((cr, -ci)).contiguous()
        if potential:
            modulus = torch.sqrt(re * re + im * im)
            phi_own = (phi[:n] - 2 * q_in * torch.log(math.pi * modulus / self.width)).contiguous()
        else:
            phi_own = torch.zeros(n, dtype=torch.float64, device=pos.device)
        self._soften(pos, q_in, pos, torch.arange(n, device=pos.device), phi_own, field,
                     potential, "Coincident charges")
        return _out(self.k_e * phi_own, like), _out(self.k_e * field, like)

    def potential(self, positions, charges, points):
        """Potential (V) at arbitrary points strictly between the plates."""
        return self.probe(positions, charges, points)[0]

    def probe(self, positions, charges, points):
        """Potential (V) and field (V/m, (2, M)) at probes (2, M) strictly between the plates
        (strip.py:241-323): no self term; the walk in the mapped plane."""
        like = _is_torch(positions) or _is_torch(points)
        pos = _points(positions, "positions")
        q_in = _device(charges).reshape(-1)
        pts = _points(points, "points")
        _, xy, q = self._sources(pos, q_in)
        pre, pim = self._plane(pts)
        probe_xy = torch.stack((pre, pim)).contiguous()
        m = probe_xy.shape[1]
        if not q.shape[0]:
            return _out(torch.zeros(m, dtype=torch.float64, device=pts.device), like), \
                _out(torch.zeros((2, m), dtype=torch.float64, device=pts.device), like)
        if self.method == "fmm":
            everything = torch.cat((xy, probe_xy), dim=1)
            tree = FlatQuadtree(xy, q, _root(everything), kernel="log", order=self.order,
                                leaf_capacity=self.leaf_capacity)
            phi, raw = tree.evaluate(probe_xy, theta=self.theta)
        else:
            phi, raw = direct_field(probe_xy, xy, q, kernel="log")
        cr, ci = self._mapped_field(raw, pre, pim)
        field = torch.stack((cr, -ci)).contiguous()
        phi = phi.contiguous().clone()
        self._soften(pos, q_in, pts, None, phi, field, True, "A probe coincides with a charge")
        return _out(self.k_e * phi, like), _out(self.k_e * field, like)
