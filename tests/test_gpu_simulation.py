"""GPU electrostatic simulation: the push kernel against the oracle's push, and the
reference's simulation tests (test_simulation.py:136-275, test_flat_tree.py:135-170,
simulation.py:120-199)."""

import sys

import numpy as np
import pytest
import sympy
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sim_module(td3):
    return sys.modules["tree_d_3d_amd.simulation"]


def two_like(td3, **kwargs):
    bodies = [td3.ChargedBody(1.0, 0.0, charge=1.0, mass=1.0),
              td3.ChargedBody(-1.0, 0.0, charge=1.0, mass=1.0)]
    return td3.ElectrostaticSimulation(bodies, td3.Quad(0.0, 0.0, 3.0), k_e=1.0, softening=0.1,
                                       dt=0.01, **kwargs)


@pytest.mark.parametrize("first_half", [True, False])
def test_push_kernel_matches_oracle(td3, oracle, sim_module, first_half):
    rng = np.random.default_rng(3)
    n = 100_003
    host = dict(positions=rng.uniform(-1, 1, (2, n)), velocities=rng.normal(0, 1, (2, n)),
                forces=np.zeros((2, n)), electric=rng.normal(0, 1, (2, n)),
                masses=rng.uniform(0.5, 2, n), charges=rng.normal(0, 1, n),
                fixed=rng.random(n) < 0.1)
    host["charges"][:7] = 0.0
    bodies = [td3.ChargedBody(0.0, 0.0) for _ in range(n)]
    sim = td3.ElectrostaticSimulation(bodies, td3.Quad(0, 0, 1), dt=0.03, B_z=1.7)
    dev = torch.device("cuda")
    sim.positions = torch.tensor(host["positions"], device=dev)
    sim.velocities = torch.tensor(host["velocities"], device=dev)
    sim._electric = torch.tensor(host["electric"], device=dev)
    sim.masses = torch.tensor(host["masses"], device=dev)
    sim.charges = torch.tensor(host["charges"], device=dev)
    sim.fixed = torch.tensor(host["fixed"], device=dev)
    sim._fixed_u8 = sim.fixed.to(torch.uint8)
    sim_module._push(sim, first_half)
    oracle.push(**host, dt=0.03, b_z=1.7, first_half=first_half)
    for name in ("positions", "velocities", "forces"):
        got = getattr(sim, name).cpu().numpy()
        np.testing.assert_allclose(got, host[name], rtol=1e-14, atol=1e-15, err_msg=name)


def test_bodies_are_views(td3):
    sim = two_like(td3)
    assert len(sim.bodies) == 2 and sim.theta == pytest.approx(0.5) and sim.B_z == 0
    assert sim.get_total_charge() == pytest.approx(2.0)
    assert sim.get_kinetic_energy() == pytest.approx(0.0)
    assert np.allclose(sim.get_center_of_mass(), [0, 0])
    assert np.allclose(sim.get_momentum(), [0, 0])
    body = sim.bodies[0]
    body.vel = (0.5, -1.0)
    assert np.array_equal(sim.velocities[:, 0].cpu().numpy(), [0.5, -1.0])
    body.add_force(0.5, 2.5)
    assert np.allclose(body.force, [0.5, 2.5])
    body.reset_force()
    assert np.allclose(body.force, [0, 0])
    single = td3.ElectrostaticSimulation(
        [td3.ChargedBody(0.0, 0.0, vx=1.0, vy=1.0, mass=2.0, charge=1.0)], td3.Quad(0, 0, 2),
        k_e=1.0)
    assert single.get_kinetic_energy() == pytest.approx(2.0)


def test_like_charges_repel_and_opposite_attract(td3):
    sim = two_like(td3)
    start = [b.pos.copy() for b in sim.bodies]
    sim.step()
    for b, p0 in zip(sim.bodies, start):
        assert not np.allclose(b.pos, p0)
    for _ in range(49):
        sim.step()
    assert np.linalg.norm(sim.bodies[1].pos - sim.bodies[0].pos) > 2.0
    bodies = [td3.ChargedBody(1.0, 0.0, charge=1.0), td3.ChargedBody(-1.0, 0.0, charge=-1.0)]
    sim = td3.ElectrostaticSimulation(bodies, td3.Quad(0, 0, 5), k_e=1.0, softening=0.1,
                                      dt=0.005)
    for _ in range(30):
        sim.step()
    assert np.linalg.norm(bodies[1].pos - bodies[0].pos) < 2.0


def test_magnetic_circular_motion(td3):
    body = td3.ChargedBody(1.0, 0.0, vx=0.0, vy=1.0)
    sim = td3.ElectrostaticSimulation([body], td3.Quad(0, 0, 5), k_e=1.0, softening=0.1,
                                      dt=0.005, B_z=1.0)
    for _ in range(200):
        sim.step()
    assert np.linalg.norm(body.vel) == pytest.approx(1.0, rel=1e-12)
    # one particle alone: its force is the Lorentz force q B_z (v_y, -v_x)
    v = body.vel
    np.testing.assert_allclose(body.force, [v[1], -v[0]], rtol=1e-14)


def test_trees_and_direct_sums(td3):
    sim = two_like(td3)
    tree = sim.build_tree()
    assert isinstance(tree, td3.FlatQuadtree)
    assert tree.root.total_charge == pytest.approx(2.0)
    assert two_like(td3, method="tree").step() is not None
    assert two_like(td3).step() is None


def test_drift_rule(td3):
    # simulation.py:120-137: one particle, no field
    vx, vy, b_z, dt = sympy.symbols("vx vy B_z dt", real=True)
    angle = b_z * dt
    rule = dt * (sympy.sin(angle) / angle * vx + (1 - sympy.cos(angle)) / angle * vy)
    rng = np.random.default_rng(8)
    for _ in range(6):
        args = dict(vx=rng.uniform(-1, 1), vy=rng.uniform(-1, 1), B_z=rng.uniform(0.1, 3),
                    dt=rng.uniform(0.1, 1))
        body = td3.ChargedBody(0.0, 0.0, args["vx"], args["vy"])
        sim = td3.ElectrostaticSimulation([body], td3.Quad(0, 0, 1), dt=args["dt"], B_z=args["B_z"])
        sim.step()
        want = float(rule.evalf(40, subs={vx: args["vx"], vy: args["vy"], b_z: args["B_z"],
                                          dt: args["dt"]}))
        assert body.pos[0] == pytest.approx(want, rel=1e-12, abs=1e-15)


def test_orbit_converges_at_second_order(td3):
    # simulation.py:140-169
    softening, b_z, time = 0.1, 0.7, 2.0
    attraction = (1 + softening ** 2) ** -1.5
    speed = 0.5 * (b_z + np.sqrt(b_z ** 2 + 4 * attraction))
    exact = np.array([np.cos(speed * time), np.sin(speed * time)])
    errors = []
    for dt in (0.04, 0.02, 0.01):
        electron = td3.ChargedBody(1.0, 0.0, vy=speed, charge=-1.0)
        ion = td3.ChargedBody(0.0, 0.0, fixed=True)
        sim = td3.ElectrostaticSimulation([electron, ion], td3.Quad(0, 0, 2), theta=0.0, k_e=1.0,
                                          softening=softening, dt=dt, B_z=b_z)
        for _ in range(round(time / dt)):
            sim.step()
        errors.append(np.linalg.norm(electron.pos - exact))
        assert np.array_equal(ion.pos, [0, 0])
    rates = np.log2(np.array(errors[:-1]) / errors[1:])
    assert np.all(rates > 1.8), (errors, rates)


@pytest.mark.parametrize("kernel", ["coulomb", "log"])
def test_pair_energy(td3, kernel):
    # simulation.py:172-199
    rng = np.random.default_rng(9)
    for _ in range(4):
        x, q1, q2, k_e = rng.uniform(0.2, 3), rng.uniform(0.5, 2), rng.uniform(0.5, 2), 1.3
        eps, r0 = rng.uniform(0, 0.5), rng.uniform(0.5, 2)
        bodies = [td3.ChargedBody(0.0, 0.0, charge=q1), td3.ChargedBody(x, 0.0, charge=q2)]
        sim = td3.ElectrostaticSimulation(bodies, td3.Quad(0, 0, 4), k_e=k_e, softening=eps,
                                          kernel=kernel, reference_radius=r0)
        if kernel == "coulomb":
            want = k_e * q1 * q2 / np.sqrt(x * x + eps * eps)
        else:
            want = -k_e * q1 * q2 * np.log((x * x + eps * eps) / r0 ** 2)
        assert sim.get_potential_energy(exact=True) == pytest.approx(want, rel=1e-13)
        assert sim.get_potential_energy() == pytest.approx(want, rel=1e-13)


def make_cloud(td3, n=1500, fixed_every=0, **kwargs):
    rng = np.random.default_rng(2)
    pos = rng.uniform(-0.9, 0.9, (2, n))
    q = rng.choice([-1.0, 1.0], n)
    bodies = [td3.ChargedBody(*pos[:, i], charge=q[i],
                              fixed=bool(fixed_every and i % fixed_every == 0))
              for i in range(n)]
    kwargs = {"k_e": 1, "softening": 0.05, "dt": 1e-3, **kwargs}
    return td3.ElectrostaticSimulation(bodies, td3.Quad(0, 0, 1), **kwargs)


@pytest.mark.parametrize("method", ["tree", "direct"])
def test_one_field_evaluation_per_step(td3, sim_module, monkeypatch, method):
    sim = make_cloud(td3, method=method)
    calls = []
    evaluate, direct = td3.FlatQuadtree.evaluate, sim_module.direct_field
    monkeypatch.setattr(td3.FlatQuadtree, "evaluate",
                        lambda *a, **k: calls.append("tree") or evaluate(*a, **k))
    monkeypatch.setattr(sim_module, "direct_field",
                        lambda *a, **k: calls.append("direct") or direct(*a, **k))
    sim.step()
    assert calls == [method, method]
    for _ in range(3):
        sim.step()
    assert calls == [method] * 5
    sim.get_potential_energy()
    assert len(calls) == 5


def test_fixed_charges_are_evaluated_on_request(td3):
    sim = make_cloud(td3, fixed_every=50, method="tree")
    sim.step()
    fixed = sim.fixed.cpu().numpy()
    moved = sim.positions.cpu().numpy()
    assert not np.any(sim.forces.cpu().numpy()[:, fixed])
    start = make_cloud(td3, fixed_every=50).positions.cpu().numpy()
    assert np.array_equal(moved[:, fixed], start[:, fixed])
    assert np.all(moved[:, ~fixed] != start[:, ~fixed])
    sim.compute_fields(include_fixed=True)
    assert np.all(sim._electric.cpu().numpy()[:, fixed] != 0)


def test_cached_forces_match_fresh_evaluation(td3):
    cached = make_cloud(td3, method="fmm", theta=0.5)
    fresh = make_cloud(td3, method="fmm", theta=0.5)
    for _ in range(5):
        cached.step()
        fresh._state = None
        fresh.step()
    assert torch.equal(cached.positions, fresh.positions)
    assert torch.equal(cached.velocities, fresh.velocities)


def test_energy_is_conserved_by_fmm_steps(td3):
    sim = make_cloud(td3, n=2000, method="fmm", theta=0.4, dt=2e-5)
    e0 = sim.get_total_energy()
    for _ in range(20):
        sim.step()
    assert abs(sim.get_total_energy() - e0) < 1e-4 * abs(e0)
    assert np.allclose(sim.get_momentum(), 0, atol=1e-8)


def test_softened_source_small_amplitude_oscillation(td3):
    softening, amplitude = 0.2, 1e-4
    omega = np.sqrt(1 / softening ** 3)
    period = 2 * np.pi / omega
    electron = td3.ChargedBody(amplitude, 0.0, charge=-1.0)
    ion = td3.ChargedBody(0.0, 0.0, charge=1.0, fixed=True)
    sim = td3.ElectrostaticSimulation([electron, ion], td3.Quad(0, 0, 2), k_e=1.0,
                                      softening=softening, dt=period / 400)
    for _ in range(400):
        sim.step()
        expected = amplitude * np.cos(omega * sim.current_time)
        assert abs(electron.pos[0] - expected) < 1e-4 * amplitude
    assert np.array_equal(ion.pos, [0, 0])
