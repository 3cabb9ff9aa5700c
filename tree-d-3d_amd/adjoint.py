"""Gradients of losses on FMM potentials, by the adjoint (transposed) FMM.

The reference's differentiable code is the smooth-membership prototype
(smooth.py:242-452): at ``band = 0`` it is the ordinary tree, its energy is
``U = 1/2 sum_i q_i phi_i`` (smooth.py:411-413) and its forces are ``-dU/dx`` checked by
central differences (tests/test_smooth.py:27-37).  It has no autograd; here the log-kernel
FMM of `FlatQuadtree` sits behind a ``torch.autograd.Function`` whose backward pass is
itself two FMM evaluations on the GPU:

    phi_t = sum_s q_s G(t, s),  G(t, s) = -ln(|t - s|^2 + eps^2)   (own charge excluded)

    dL/dq_s   = psi(s),         psi(s) = sum_t g_t G(t, s)
    dL/dx_s   = -g_s E_s - q_s E^psi(s)                            (E^psi = -grad psi)
    dL/dp     = -g_p E_p                                           (probe positions)

with ``g = dL/dphi`` at every target (the sources themselves and the probes).  ``psi`` is
the potential of charges ``g`` placed at the targets, so the backward pass re-runs the
source tree's FMM with the charges ``g_src`` (`set_charges`: the tree and its interaction
lists are kept, the expansions are recomputed) and runs a probe tree with charges
``g_probe`` at the sources (``fmm_at``).  The near field is recomputed, not stored.

A loss on the fields ``E_t = -grad_t phi_t`` as well, with ``h = dL/dE`` at every target,
adds a dipole ``h_t`` at each target to those charges (``sum_t h_t . dE_t/dq_s`` is the
potential at ``s`` of dipoles ``h``: ``2 h . (t - s) / |t - s|^2``) and the field's
Jacobian ``J = dE/dx`` (the second-derivative kernel) at the targets themselves:

    dL/dq_s   = psi(s)                    (charges g and dipoles h at the targets)
    dL/dx_s   = -g_s E_s + J_s h_s - q_s E^psi(s)
    dL/dp     = -g_p E_p + J_p h_p
"""

from __future__ import annotations

import torch

from . import _abi
from .flat_tree import FlatQuadtree
from .quad import Quad


def _covering_root(points: torch.Tensor) -> Quad:
    lo = points.min(dim=1).values
    hi = points.max(dim=1).values
    center = 0.5 * lo + 0.5 * hi
    half = float(0.5 * (hi - lo).max()) * (1 + 1e-12) + 1e-300
    return Quad(float(center[0]), float(center[1]), half)


class _LogPotentials(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, charges, probes, root, options):
        theta, softening = options["theta"], options["softening"]
        tree_kw = _tree_kw(options)
        pos = positions.detach().contiguous()
        q = charges.detach().contiguous()
        tree = FlatQuadtree(pos, q, root, **tree_kw)
        phi, field = tree.fmm(theta=theta, softening=softening)
        pr = probes.detach().contiguous()
        if pr.shape[1]:
            phi_p, field_p = tree.fmm_at(pr, theta=theta, softening=softening)
        else:
            phi_p, field_p = pr.new_zeros(0), pr.new_zeros((2, 0))
        ctx.save_for_backward(pos, q, pr, field, field_p)
        ctx.tree, ctx.options, ctx.tree_kw = tree, options, tree_kw
        return phi, phi_p

    @staticmethod
    def backward(ctx, g_src, g_probe):
        pos, q, pr, field, field_p = ctx.saved_tensors
        g_src = _zeros_if_none(g_src, (pos.shape[1],), pos)
        psi, e_psi = _transposed(ctx, pos, pr, g_src, None, g_probe, None)
        grad_pos = -g_src[None, :] * field - q[None, :] * e_psi
        grad_probes = None
        if pr.shape[1] and ctx.needs_input_grad[2]:
            gp = _zeros_if_none(g_probe, (pr.shape[1],), pr)
            grad_probes = -gp[None, :] * field_p
        return grad_pos, psi, grad_probes, None, None


class _LogFields(torch.autograd.Function):
    @staticmethod
    def forward(ctx, positions, charges, probes, root, options):
        theta, softening = options["theta"], options["softening"]
        tree_kw = _tree_kw(options)
        pos = positions.detach().contiguous()
        q = charges.detach().contiguous()
        tree = FlatQuadtree(pos, q, root, **tree_kw)
        phi, field, jac = tree.fmm(theta=theta, softening=softening, jacobian=True)
        pr = probes.detach().contiguous()
        if pr.shape[1]:
            phi_p, field_p, jac_p = tree.fmm_at(pr, theta=theta, softening=softening,
                                                jacobian=True)
        else:
            phi_p, field_p, jac_p = pr.new_zeros(0), pr.new_zeros((2, 0)), pr.new_zeros((3, 0))
        ctx.save_for_backward(pos, q, pr, field, field_p, jac, jac_p)
        ctx.tree, ctx.options, ctx.tree_kw = tree, options, tree_kw
        return phi, field, phi_p, field_p

    @staticmethod
    def backward(ctx, g_src, h_src, g_probe, h_probe):
        pos, q, pr, field, field_p, jac, jac_p = ctx.saved_tensors
        g_src = _zeros_if_none(g_src, (pos.shape[1],), pos)
        psi, e_psi = _transposed(ctx, pos, pr, g_src, h_src, g_probe, h_probe)
        grad_pos = -g_src[None, :] * field - q[None, :] * e_psi
        if h_src is not None:
            grad_pos = grad_pos + _jacobian_times(jac, h_src.to(torch.float64))
        grad_probes = None
        if pr.shape[1] and ctx.needs_input_grad[2]:
            gp = _zeros_if_none(g_probe, (pr.shape[1],), pr)
            grad_probes = -gp[None, :] * field_p
            if h_probe is not None:
                grad_probes = grad_probes + _jacobian_times(jac_p, h_probe.to(torch.float64))
        return grad_pos, psi, grad_probes, None, None


def _tree_kw(options) -> dict:
    return dict(kernel="log", order=options["order"], leaf_capacity=options["leaf_capacity"],
                precision=options["precision"])


def _zeros_if_none(g, shape, like):
    if g is None:
        return torch.zeros(shape, dtype=torch.float64, device=like.device)
    return g.to(torch.float64).contiguous()


def _nonzero(t) -> bool:
    return t is not None and t.numel() > 0 and bool((t != 0).any())


def _jacobian_times(jac, h):
    """``J h`` for the symmetric Jacobians ``(3, M)`` (xx, xy, yy) and vectors ``(2, M)``."""
    return torch.stack((jac[0] * h[0] + jac[1] * h[1], jac[1] * h[0] + jac[2] * h[1]))


def _transposed(ctx, pos, pr, g_src, h_src, g_probe, h_probe):
    """The transposed FMM: potential and field at the sources (own terms excluded) of
    charges ``g`` and dipoles ``h`` at every target (the sources and the probes)."""
    theta, softening = ctx.options["theta"], ctx.options["softening"]
    tree = ctx.tree
    dip = h_src.to(torch.float64).contiguous() if _nonzero(h_src) else None
    tree.set_charges(g_src, dipoles=dip)
    psi, e_psi = tree.fmm(theta=theta, softening=softening)
    if pr.shape[1] and (_nonzero(g_probe) or _nonzero(h_probe)):
        # ... and from the charges and dipoles at the probes
        gp = _zeros_if_none(g_probe, (pr.shape[1],), pr)
        probe_tree = FlatQuadtree(pr, gp, _covering_root(pr), **ctx.tree_kw)
        if _nonzero(h_probe):
            probe_tree.set_charges(gp, dipoles=h_probe.to(torch.float64).contiguous())
        psi_p, e_psi_p = probe_tree.fmm_at(pos, theta=theta, softening=softening)
        psi = psi + psi_p
        e_psi = e_psi + e_psi_p
    return psi, e_psi


def log_potentials(positions, charges, probes=None, *, root=None, order: int = 12,
                   theta: float = 0.5, softening: float = 0.0, leaf_capacity: int = 8,
                   precision: str = "fp64"):
    """The log-kernel FMM potential at every source (its own charge excluded) and at the
    ``probes`` ``(2, M)``, differentiable in ``positions`` ``(2, N)``, ``charges`` ``(N,)``
    and ``probes``: the raw potential ``-sum q ln(r^2 + eps^2)`` of `FlatQuadtree.fmm`
    (softened near field).  ``root`` defaults to a square covering the sources."""
    args = _prepare(positions, charges, probes, root, order, theta, softening, leaf_capacity,
                    precision)
    return _LogPotentials.apply(*args)


def _prepare(positions, charges, probes, root, order, theta, softening, leaf_capacity,
             precision):
    dev = _abi.device()
    positions = torch.as_tensor(positions, dtype=torch.float64, device=dev)
    charges = torch.as_tensor(charges, dtype=torch.float64, device=dev)
    if probes is None:
        probes = torch.zeros((2, 0), dtype=torch.float64, device=dev)
    probes = torch.as_tensor(probes, dtype=torch.float64, device=dev)
    if positions.ndim != 2 or positions.shape[0] != 2 or probes.ndim != 2 or probes.shape[0] != 2:
        raise ValueError("positions and probes must have shape (2, N) (component-major)")
    if charges.shape != (positions.shape[1],):
        raise ValueError("charges must have one entry per position")
    if root is None:
        root = _covering_root(positions.detach())
    options = dict(order=int(order), theta=float(theta), softening=float(softening),
                   leaf_capacity=int(leaf_capacity), precision=precision)
    return positions, charges, probes, root, options


def log_fields(positions, charges, probes=None, *, root=None, order: int = 12,
               theta: float = 0.5, softening: float = 0.0, leaf_capacity: int = 8,
               precision: str = "fp64"):
    """`log_potentials` with the fields: ``(phi (N,), E (2, N), phi_p (M,), E_p (2, M))`` at
    the sources (own charge excluded) and the probes, all differentiable in ``positions``,
    ``charges`` and ``probes``, so a loss may use potentials and fields.  The forward pass
    also keeps the field's Jacobian at every target for the backward pass."""
    args = _prepare(positions, charges, probes, root, order, theta, softening, leaf_capacity,
                    precision)
    return _LogFields.apply(*args)


def log_energy(positions, charges, **options) -> torch.Tensor:
    """``U = 1/2 sum_i q_i phi_i`` (smooth.py:411-413 at band 0; the units of
    `FlatQuadtree`: pair energy ``-q_i q_j ln(r^2 + eps^2)``), differentiable."""
    phi, _ = log_potentials(positions, charges, **options)
    charges = torch.as_tensor(charges, dtype=torch.float64, device=phi.device)
    return 0.5 * torch.dot(charges, phi)
