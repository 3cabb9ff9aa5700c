"""Adjoint gradients of losses on FMM potentials (log kernel) against FP64 torch autograd
through a direct sum and against central finite differences of the FMM loss itself
(the pattern of tests/test_smooth.py:27-37 in the reference)."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def uniform(n, m, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1, 1, (2, n))
    q = rng.choice([-1.0, 1.0], n)
    probes = rng.uniform(-1, 1, (2, m))
    return pos, q, probes


def torch_direct(pos, q, probes, eps):
    """phi at the sources (own charge excluded) and at the probes, FP64, differentiable."""
    d2 = ((pos[:, :, None] - pos[:, None, :]) ** 2).sum(0) + eps * eps
    n = pos.shape[1]
    eye = torch.eye(n, dtype=torch.float64, device=pos.device)
    kern = -torch.log(d2 + eye) * (1 - eye)
    phi = kern @ q
    dp2 = ((probes[:, :, None] - pos[:, None, :]) ** 2).sum(0) + eps * eps
    phi_p = -torch.log(dp2) @ q
    return phi, phi_p


def loss_of(phi, phi_p, q):
    # config 5: mean squared potential at the probes plus the energy 1/2 sum q phi
    return (phi_p ** 2).mean() + 0.5 * torch.dot(q, phi)


def relative(a, b):
    return float((a - b).abs().max() / b.abs().max())


def test_adjoint_matches_torch_autograd_of_direct_sum(td3):
    # unsoftened: the FMM softens only its near pairs, so only eps = 0 has a direct-sum
    # equivalent; the softened case is checked by finite differences below
    eps = 0.0
    pos_h, q_h, probes_h = uniform(8192, 512, 3)
    dev = torch.device("cuda")
    pos = torch.tensor(pos_h, device=dev, requires_grad=True)
    q = torch.tensor(q_h, device=dev, requires_grad=True)
    probes = torch.tensor(probes_h, device=dev, requires_grad=True)
    print(f"adjoint vs torch eps={eps}", flush=True)
    phi, phi_p = td3.log_potentials(pos, q, probes, root=td3.Quad(0, 0, 1), softening=eps)
    got = torch.autograd.grad(loss_of(phi, phi_p, q), (pos, q, probes))
    rphi, rphi_p = torch_direct(pos, q, probes, eps)
    want = torch.autograd.grad(loss_of(rphi, rphi_p, q), (pos, q, probes))
    # the FMM's truncation error at order 12, and the 1e-5 of the gradient target
    assert relative(phi.detach(), rphi.detach()) < 1e-7
    for g, w in zip(got, want):
        assert relative(g, w) < 1e-5


def test_adjoint_matches_central_differences(td3):
    # h = 1e-6, random particles, the FMM loss itself
    pos_h, q_h, probes_h = uniform(20000, 1024, 4)
    dev = torch.device("cuda")
    root = td3.Quad(0, 0, 1)
    kw = dict(root=root, softening=1e-3)

    def loss(pos, q):
        phi, phi_p = td3.log_potentials(pos, q, torch.tensor(probes_h, device=dev), **kw)
        return loss_of(phi, phi_p, q)

    pos = torch.tensor(pos_h, device=dev, requires_grad=True)
    q = torch.tensor(q_h, device=dev, requires_grad=True)
    g_pos, g_q = torch.autograd.grad(loss(pos, q), (pos, q))
    scale_pos, scale_q = float(g_pos.abs().max()), float(g_q.abs().max())
    h = 1e-6
    for i in np.random.default_rng(5).choice(20000, 8, replace=False):
        for axis in range(2):
            pp, pm = pos_h.copy(), pos_h.copy()
            pp[axis, i] += h
            pm[axis, i] -= h
            fd = (float(loss(torch.tensor(pp, device=dev), torch.tensor(q_h, device=dev)))
                  - float(loss(torch.tensor(pm, device=dev), torch.tensor(q_h, device=dev)))) / (2 * h)
            assert abs(fd - float(g_pos[axis, i])) < 1e-5 * scale_pos
        qp, qm = q_h.copy(), q_h.copy()
        qp[i] += h
        qm[i] -= h
        fd = (float(loss(torch.tensor(pos_h, device=dev), torch.tensor(qp, device=dev)))
              - float(loss(torch.tensor(pos_h, device=dev), torch.tensor(qm, device=dev)))) / (2 * h)
        assert abs(fd - float(g_q[i])) < 1e-5 * scale_q


def test_energy_gradient_is_minus_the_force(td3):
    pos_h, q_h, _ = uniform(3000, 0, 6)
    dev = torch.device("cuda")
    pos = torch.tensor(pos_h, device=dev, requires_grad=True)
    q = torch.tensor(q_h, device=dev)
    energy = td3.log_energy(pos, q, root=td3.Quad(0, 0, 1), softening=1e-3)
    (g,) = torch.autograd.grad(energy, (pos,))
    tree = td3.FlatQuadtree(pos_h, q_h, td3.Quad(0, 0, 1), kernel="log")
    _, field = tree.fmm(softening=1e-3)
    force = q_h * field
    np.testing.assert_allclose(-g.cpu().numpy(), force, rtol=0, atol=1e-10 * np.abs(force).max())


def test_symmetric_fmm_forces_are_the_energy_gradient(td3):
    # smooth.py:411-422 at band 0 and tests/test_smooth.py:27-37: with the symmetric
    # truncation (k + l <= p) the tree's forces are -dU/dx, U = 1/2 sum q phi, to central
    # differences (h = 1e-6) within 1e-7 max|F|
    rng = np.random.default_rng(7)
    pos = rng.uniform(-0.9, 0.9, (2, 3000))
    q = rng.choice([-1.0, 1.0], 3000)
    kw = dict(kernel="log", order=12, symmetric=True, leaf_capacity=8)

    def energy(p):
        tree = td3.FlatQuadtree(p, q, td3.Quad(0, 0, 1), **kw)
        phi, _ = tree.fmm(theta=0.5, softening=0.02)
        return 0.5 * float(np.dot(q, phi))

    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), **kw)
    _, field = tree.fmm(theta=0.5, softening=0.02)
    force = q * field
    scale = np.abs(force).max()
    h = 1e-6
    for i in rng.choice(3000, 6, replace=False):
        for axis in range(2):
            pp, pm = pos.copy(), pos.copy()
            pp[axis, i] += h
            pm[axis, i] -= h
            fd = -(energy(pp) - energy(pm)) / (2 * h)
            assert abs(fd - force[axis, i]) < 1e-7 * scale, (i, axis, fd, force[axis, i])


# ------------------------------------------------------------------ losses on fields
def torch_fields(pos, q, targets, eps, own):
    """Potential and field at ``targets`` of the charges ``q`` at ``pos`` (FP64 direct sum,
    differentiable); ``own``: the targets are the sources, each excluding itself."""
    d = targets[:, :, None] - pos[:, None, :]
    r2 = d[0] ** 2 + d[1] ** 2 + eps * eps
    keep = 1.0
    if own:
        # the own pair is masked out (r2 = 1 + eps^2 there keeps it finite)
        eye = torch.eye(pos.shape[1], dtype=torch.float64, device=pos.device)
        r2 = r2 + eye
        keep = 1.0 - eye
    return -((torch.log(r2) * keep) @ q), 2 * ((d / r2) * (q * keep)).sum(2)


def test_dipole_sources_match_direct_sum(td3):
    # a dipole D at s adds -2 D.d / r^2 to the potential at t (d = t - s)
    rng = np.random.default_rng(21)
    n = 6000
    pos = rng.uniform(-1, 1, (2, n))
    q = rng.choice([-1.0, 1.0], n)
    dip = rng.normal(0, 1, (2, n))
    probes = rng.uniform(-1, 1, (2, 700))
    dev = torch.device("cuda")
    print("dipole sources vs direct", flush=True)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", order=16)
    tree.set_charges(q, dipoles=dip)
    phi, field = tree.fmm(theta=0.5)
    phi_p, field_p = tree.fmm_at(probes, theta=0.5)
    x, c, D = (torch.tensor(a, device=dev) for a in (pos, q, dip))
    for targets, own, got in ((x, True, (phi, field)), (torch.tensor(probes, device=dev), False,
                                                        (phi_p, field_p))):
        d = targets[:, :, None] - x[:, None, :]
        r2 = d[0] ** 2 + d[1] ** 2
        if own:
            r2 = r2 + torch.eye(n, dtype=torch.float64, device=dev)
        pd = (D[:, None, :] * d).sum(0)
        rphi = -(torch.log(r2) @ c) - (2 * pd / r2).sum(1)
        rfield = 2 * ((d / r2) * c).sum(2) + (2 * D[:, None, :] / r2 - 4 * pd * d / r2 ** 2).sum(2)
        # a dipole's expansion converges one order slower than a charge's
        assert relative(torch.as_tensor(got[0], device=dev), rphi) < 1e-7
        assert relative(torch.as_tensor(got[1], device=dev), rfield) < 1e-6
    # set_charges without dipoles drops them
    tree.set_charges(q)
    phi0, _ = tree.fmm(theta=0.5)
    rphi0, _ = torch_fields(x, c, x, 0.0, True)
    assert relative(torch.as_tensor(phi0, device=dev), rphi0) < 1e-8


def test_field_jacobian_matches_autograd_of_direct_sum(td3):
    # J = dE/dx at a target, its own position moved alone: autograd of the direct field at
    # copies of the positions, each excluding its own index (unsoftened: the FMM softens
    # only its near pairs, so only eps = 0 has a direct-sum equivalent)
    rng = np.random.default_rng(22)
    n, m = 6000, 700
    pos = rng.uniform(-1, 1, (2, n))
    q = rng.choice([-1.0, 1.0], n)
    probes = rng.uniform(-1, 1, (2, m))
    dev = torch.device("cuda")
    print("field jacobian vs autograd", flush=True)
    x, c = torch.tensor(pos, device=dev), torch.tensor(q, device=dev)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", order=14)
    _, _, jac = tree.fmm(theta=0.5, jacobian=True)
    _, _, jac_p = tree.fmm_at(probes, theta=0.5, jacobian=True)
    for pts, own, got in ((pos, True, jac), (probes, False, jac_p)):
        t = torch.tensor(pts, device=dev, requires_grad=True)
        _, e = torch_fields(x, c, t, 0.0, own)
        gx = torch.autograd.grad(e[0].sum(), t, retain_graph=True)[0]
        gy = torch.autograd.grad(e[1].sum(), t)[0]
        want = torch.stack((gx[0], gx[1], gy[1]))
        assert relative(torch.as_tensor(got, device=dev), want) < 1e-6
        assert relative(gx[1], gy[0]) < 1e-12  # the Jacobian is symmetric


def test_mixed_fmm_with_jacobian_and_dipoles(td3):
    """A fresh tree's first fmm in mixed precision with jacobian=True, and with dipoles,
    takes rows that were read back (not the deferred rows of the plain mixed FMM): it
    returns, records fmm_pairs, and stays within 1e-6 of the FP64 results."""
    rng = np.random.default_rng(31)
    n = 5000
    pos = rng.uniform(-1, 1, (2, n))
    q = rng.choice([-1.0, 1.0], n)
    dip = rng.normal(0, 1, (2, n))
    print("mixed fmm jacobian and dipoles", flush=True)
    fp64 = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", order=12)
    want = fp64.fmm(theta=0.5, softening=1e-3, jacobian=True)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", order=12, precision="mixed")
    got = tree.fmm(theta=0.5, softening=1e-3, jacobian=True)
    assert tree.fmm_pairs == fp64.fmm_pairs
    for a, b in zip(got, want):
        assert relative(torch.as_tensor(a), torch.as_tensor(b)) < 1e-6
    fp64.set_charges(q, dipoles=dip)
    want = fp64.fmm(theta=0.5, softening=1e-3)
    dtree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", order=12,
                             precision="mixed")
    dtree.set_charges(q, dipoles=dip)
    got = dtree.fmm(theta=0.5, softening=1e-3)
    assert dtree.fmm_pairs == fp64.fmm_pairs
    for a, b in zip(got, want):
        assert relative(torch.as_tensor(a), torch.as_tensor(b)) < 1e-6


def test_softened_near_field_jacobian_matches_autograd(td3):
    # with one leaf every pair is a near (softened) pair, so the softened direct sum is
    # the FMM's own semantics: the near-field Jacobian 2 q (I / rho - 2 d d^T / rho^2)
    rng = np.random.default_rng(28)
    n, m, eps = 3000, 500, 0.01
    pos = rng.uniform(-1, 1, (2, n))
    q = rng.choice([-1.0, 1.0], n)
    probes = rng.uniform(-1, 1, (2, m))
    dev = torch.device("cuda")
    print("softened near-field jacobian vs autograd", flush=True)
    x, c = torch.tensor(pos, device=dev), torch.tensor(q, device=dev)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", order=8,
                            leaf_capacity=n)
    _, _, jac = tree.fmm(theta=0.5, softening=eps, jacobian=True)
    _, _, jac_p = tree.fmm_at(probes, theta=0.5, softening=eps, jacobian=True,
                              leaf_capacity=m)
    for pts, own, got in ((pos, True, jac), (probes, False, jac_p)):
        t = torch.tensor(pts, device=dev, requires_grad=True)
        _, e = torch_fields(x, c, t, eps, own)
        gx = torch.autograd.grad(e[0].sum(), t, retain_graph=True)[0]
        gy = torch.autograd.grad(e[1].sum(), t)[0]
        want = torch.stack((gx[0], gx[1], gy[1]))
        assert relative(torch.as_tensor(got, device=dev), want) < 1e-12


def field_loss(phi, field, phi_p, field_p, q, w):
    # potentials and fields at the sources and probes: energy, a weighted field term,
    # mean squared probe potential and field
    return (0.5 * torch.dot(q, phi) + (w * field).sum() + (phi_p ** 2).mean()
            + (field_p ** 2).sum(0).mean())


def test_field_loss_adjoint_matches_torch_autograd_of_direct_sum(td3):
    pos_h, q_h, probes_h = uniform(8192, 512, 23)
    w_h = np.random.default_rng(24).normal(0, 1e-3, (2, 8192))
    dev = torch.device("cuda")
    pos = torch.tensor(pos_h, device=dev, requires_grad=True)
    q = torch.tensor(q_h, device=dev, requires_grad=True)
    probes = torch.tensor(probes_h, device=dev, requires_grad=True)
    w = torch.tensor(w_h, device=dev)
    print("field-loss adjoint vs torch", flush=True)
    out = td3.log_fields(pos, q, probes, root=td3.Quad(0, 0, 1), order=14)
    got = torch.autograd.grad(field_loss(*out, q, w), (pos, q, probes))
    phi, field = torch_fields(pos, q, pos, 0.0, True)
    phi_p, field_p = torch_fields(pos, q, probes, 0.0, False)
    assert relative(out[1].detach(), field.detach()) < 1e-7
    want = torch.autograd.grad(field_loss(phi, field, phi_p, field_p, q, w), (pos, q, probes))
    for g, h in zip(got, want):
        assert relative(g, h) < 1e-5


def test_field_loss_adjoint_matches_central_differences(td3):
    pos_h, q_h, probes_h = uniform(20000, 1024, 25)
    w_h = np.random.default_rng(26).normal(0, 1e-3, (2, 20000))
    dev = torch.device("cuda")
    w = torch.tensor(w_h, device=dev)
    kw = dict(root=td3.Quad(0, 0, 1), softening=1e-3)

    def loss(p, c):
        out = td3.log_fields(p, c, torch.tensor(probes_h, device=dev), **kw)
        return field_loss(*out, c, w)

    pos = torch.tensor(pos_h, device=dev, requires_grad=True)
    q = torch.tensor(q_h, device=dev, requires_grad=True)
    print("field-loss adjoint vs central differences", flush=True)
    g_pos, g_q = torch.autograd.grad(loss(pos, q), (pos, q))
    scale_pos, scale_q = float(g_pos.abs().max()), float(g_q.abs().max())
    h = 1e-6
    qt = torch.tensor(q_h, device=dev)
    for i in np.random.default_rng(27).choice(20000, 6, replace=False):
        for axis in range(2):
            pp, pm = pos_h.copy(), pos_h.copy()
            pp[axis, i] += h
            pm[axis, i] -= h
            fd = (float(loss(torch.tensor(pp, device=dev), qt))
                  - float(loss(torch.tensor(pm, device=dev), qt))) / (2 * h)
            assert abs(fd - float(g_pos[axis, i])) < 1e-5 * scale_pos, (i, axis)
        qp, qm = q_h.copy(), q_h.copy()
        qp[i] += h
        qm[i] -= h
        pt = torch.tensor(pos_h, device=dev)
        fd = (float(loss(pt, torch.tensor(qp, device=dev)))
              - float(loss(pt, torch.tensor(qm, device=dev)))) / (2 * h)
        assert abs(fd - float(g_q[i])) < 1e-5 * scale_q, i
