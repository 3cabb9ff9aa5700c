"""The batched ensemble: every system's forest tree and walk equal its own tree's, and
ensemble steps equal the systems' own ElectrostaticSimulation steps (method "tree")."""

import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def gaussian_systems(td3, sizes, seed, **kwargs):
    rng = np.random.default_rng(seed)
    sims = []
    for k in sizes:
        pos = rng.normal(0.0, 0.1, (2, k)).clip(-0.95, 0.95)
        bodies = [td3.ChargedBody(pos[0, i], pos[1, i], charge=1.0, mass=1.0) for i in range(k)]
        kw = dict(theta=0.5, k_e=1.0, softening=1e-3, dt=1e-4, B_z=1.0, method="tree")
        kw.update(kwargs)
        sims.append(td3.ElectrostaticSimulation(bodies, td3.Quad(0.0, 0.0, 1.0), **kw))
    return sims


def test_forest_walk_equals_each_systems_own_tree(td3):
    ens_mod = sys.modules["tree_d_3d_amd.ensemble"]
    rng = np.random.default_rng(2)
    sizes = [1, 9, 300, 4097, 12000, 64, 2500]
    roots = [(0.0, 0.0, 1.0), (0.2, -0.1, 1.5), (0.0, 0.0, 1.0), (0.01, 0.02, 0.99),
             (0.0, 0.0, 2.0), (-0.3, 0.3, 1.7), (0.0, 0.0, 1.0)]
    parts = [rng.normal(0, 0.12, (2, k)).clip(-0.95, 0.95) for k in sizes]
    pos = torch.tensor(np.concatenate(parts, axis=1), device="cuda")
    q = torch.tensor(rng.choice([-1.0, 1.0], pos.shape[1]), device="cuda")
    start = torch.tensor(np.concatenate(([0], np.cumsum(sizes))), device="cuda")
    system = torch.repeat_interleave(torch.arange(len(sizes), dtype=torch.int32, device="cuda"),
                                     torch.tensor(sizes, device="cuda"))
    print("forest walk", flush=True)
    forest = ens_mod._Forest(pos, q, system, start,
                             torch.tensor(roots, dtype=torch.float64, device="cuda"), 8, 48)
    phi, field = forest.walk(0.5, 1e-3)
    total = len(sizes)
    for s, (k, root) in enumerate(zip(sizes, roots)):
        sl = slice(int(start[s]), int(start[s + 1]))
        single = td3.FlatQuadtree(pos[:, sl], q[sl], td3.Quad(*root), leaf_capacity=8)
        total += single.node_count - 1
        want_phi, want_field = single.evaluate(pos[:, sl], exclude=np.arange(k), theta=0.5,
                                               softening=1e-3)
        assert torch.equal(phi[sl], want_phi) and torch.equal(field[:, sl], want_field)
    assert forest.node_count == total


def test_walk_counts_closed_forms(td3):
    """td3_walk_counts: the counted walk's fields equal the plain walk's bit for bit; at
    theta 0 no cell is accepted and every target meets every other particle of its system
    (sum of n_s (n_s - 1) terms, and each wavefront reads all of its systems' sources at
    least once); at theta 0.5 cells are accepted and fewer terms are summed."""
    ens_mod = sys.modules["tree_d_3d_amd.ensemble"]
    rng = np.random.default_rng(12)
    sizes = [1, 70, 640, 2000]
    pos = torch.tensor(np.concatenate([rng.normal(0, 0.12, (2, k)).clip(-0.95, 0.95)
                                       for k in sizes], axis=1), device="cuda")
    q = torch.tensor(rng.choice([-1.0, 1.0], pos.shape[1]), device="cuda")
    start = torch.tensor(np.concatenate(([0], np.cumsum(sizes))), device="cuda")
    system = torch.repeat_interleave(torch.arange(len(sizes), dtype=torch.int32, device="cuda"),
                                     torch.tensor(sizes, device="cuda"))
    roots = torch.tensor([(0.0, 0.0, 1.0)] * len(sizes), dtype=torch.float64, device="cuda")
    forest = ens_mod._Forest(pos, q, system, start, roots, 8, 48)
    exact = forest.walk_counts(0.0, 1e-3)
    print("theta 0 counts", exact, flush=True)
    assert exact["accepted_per_target"] == 0 and exact["accepted_per_wave"] == 0
    assert exact["terms_per_target"] == sum(k * (k - 1) for k in sizes)
    assert exact["sources_per_wave"] >= sum(sizes)
    far = forest.walk_counts(0.5, 1e-3)
    assert far["accepted_per_target"] > 0
    assert far["accepted_per_wave"] <= far["accepted_per_target"]
    assert far["terms_per_target"] < exact["terms_per_target"]
    phi, field = forest.walk(0.5, 1e-3)
    # the counted launch writes the same fields
    from tree_d_3d_amd import _abi
    d = forest._d
    n = pos.shape[1]
    phi2 = torch.zeros(n, dtype=torch.float64, device="cuda")
    field2 = torch.zeros((2, n), dtype=torch.float64, device="cuda")
    counts = torch.zeros(4, dtype=torch.int64, device="cuda")
    status = torch.zeros(1, dtype=torch.int64, device="cuda")
    _abi.call("td3_walk_counts", _abi.COULOMB, _abi.FP64, _abi.ptr(pos), n, _abi.ptr(d["perm"]),
              _abi.ptr(torch.arange(n, device="cuda")), _abi.ptr(forest._sorted_pos),
              _abi.ptr(forest._scharge), _abi.ptr(d["perm"]), n, _abi.ptr(d["node_center"]),
              _abi.ptr(d["node_half"]), _abi.ptr(forest._node_info), _abi.ptr(d["node_qabs"]),
              _abi.ptr(d["node_q"]), _abi.ptr(d["node_p"]), _abi.ptr(d["node_s"]),
              _abi.ptr(d["moment_re"]), _abi.ptr(d["moment_im"]), forest.order, forest._depth,
              _abi.ptr(forest._system), 0.5, 1e-3, 1, _abi.ptr(phi2), _abi.ptr(field2),
              _abi.ptr(counts), _abi.ptr(status), _abi.stream())
    assert int(status.item()) == 0
    assert torch.equal(phi, phi2) and torch.equal(field, field2)


def test_ensemble_steps_equal_each_simulation(td3):
    sizes = [1500, 3000, 700, 5000]
    alone = gaussian_systems(td3, sizes, seed=4)
    together = td3.Ensemble(gaussian_systems(td3, sizes, seed=4))
    print("ensemble steps", flush=True)
    for _ in range(5):
        for sim in alone:
            sim.step()
        together.step()
    for s, sim in enumerate(alone):
        sl = together.system_slice(s)
        np.testing.assert_allclose(together.positions[:, sl].cpu().numpy(),
                                   sim.positions.cpu().numpy(), rtol=0, atol=1e-14)
        np.testing.assert_allclose(together.velocities[:, sl].cpu().numpy(),
                                   sim.velocities.cpu().numpy(), rtol=1e-12, atol=1e-12)
    energies = together.total_energies()
    want = np.array([sim.get_total_energy() for sim in alone])
    np.testing.assert_allclose(energies, want, rtol=1e-12)


def test_ensemble_roots_grow_and_energy_drift_is_small(td3):
    # particles pushed out of their root square: each system's root grows on its own
    sims = gaussian_systems(td3, [800, 2000], seed=6, dt=1e-5)
    sims[1].velocities[0] += 3000.0
    ens = td3.Ensemble(sims)
    e0 = ens.total_energies()
    for _ in range(40):
        ens.step()
    roots = ens.roots.cpu().numpy()
    assert roots[0, 2] == 1.0 and roots[1, 2] > 1.0
    drift = np.abs(ens.total_energies() - e0) / np.abs(e0)
    assert drift[0] < 1e-3 and drift[1] < 1e-3, drift


def test_autotuned_walk_shapes_give_the_same_fields(td3):
    # every launch shape of the walk computes the same walk (a wavefront's targets and
    # traversal do not depend on the wavefronts per block): bit-identical fields
    abi = sys.modules["tree_d_3d_amd._abi"]
    sims = gaussian_systems(td3, [2000, 700, 5000, 33], 9)
    ens = td3.Ensemble(sims)
    print("walk launch shapes", flush=True)
    previous = abi.lib().td3_walk_waves(0)
    try:
        fields = []
        for waves in (1, 2, 4, 8):
            abi.lib().td3_walk_waves(waves)
            phi, field = ens.compute_fields().walk(ens.theta, ens.softening)
            fields.append((phi.clone(), field.clone()))
        for phi, field in fields[1:]:
            assert torch.equal(phi, fields[0][0]) and torch.equal(field, fields[0][1])
        chosen = ens.autotune(repeats=1)
        assert chosen in (1, 2, 4, 8) and abi.lib().td3_walk_waves(0) == chosen
    finally:
        abi.lib().td3_walk_waves(previous)


def _same_state(a, b):
    for name in ("positions", "velocities", "forces"):
        x, y = getattr(a, name), getattr(b, name)
        assert torch.equal(x, y), (name, float((x - y).abs().max()))
    a.compute_fields()
    b.compute_fields()
    assert torch.equal(a._potential, b._potential) and torch.equal(a._electric, b._electric)
    # per-system sums in a fixed order (td3_system_energies): equal states, equal energies
    assert np.array_equal(a.total_energies(), b.total_energies())


def test_system_energies_equal_fp64_sums(td3):
    """td3_system_energies (a workgroup per system, fixed order) against NumPy's FP64 sums
    of 1/2 m |v|^2 and 1/2 q phi per system, and bit-equal from call to call."""
    sizes = [1, 700, 3000, 255, 257]
    ens = td3.Ensemble(gaussian_systems(td3, sizes, seed=6))
    print("system energies", flush=True)
    ens.run(3, graph=False)
    kin, pot = ens.kinetic_energies(), ens.potential_energies()
    m, q = ens.masses.cpu().numpy(), ens.charges.cpu().numpy()
    v, phi = ens.velocities.cpu().numpy(), ens._potential.cpu().numpy()
    start = np.concatenate(([0], np.cumsum(sizes)))
    for s in range(len(sizes)):
        sl = slice(start[s], start[s + 1])
        want_k = np.sum(0.5 * m[sl] * (v[0, sl] ** 2 + v[1, sl] ** 2))
        want_p = np.sum(0.5 * q[sl] * phi[sl])
        assert abs(kin[s] - want_k) <= 1e-13 * abs(want_k) + 1e-300
        assert abs(pot[s] - want_p) <= 1e-13 * abs(want_p)
    assert np.array_equal(kin, ens.kinetic_energies())
    assert np.array_equal(pot, ens.potential_energies())


def test_graph_replays_equal_steps(td3):
    """Ensemble.run(graph=True) (one step direct, then a recorded HIP graph replayed) gives
    the states of Ensemble.step, bit for bit: after a first run, after every single replay
    with the other ensemble's eager steps in between, and after a long run."""
    rng = np.random.default_rng(12)
    sizes = [int(k) for k in np.exp(rng.uniform(np.log(200), np.log(6000), 24)).round()]
    eager = td3.Ensemble(gaussian_systems(td3, sizes, seed=12))
    graph = td3.Ensemble(gaussian_systems(td3, sizes, seed=12))
    print(f"graph run: {len(sizes)} systems, N = {sum(sizes)}", flush=True)
    graph.run(3, graph=True)
    for _ in range(3):
        eager.step()
    _same_state(eager, graph)
    assert graph._graph is not None and graph.iteration == 3
    for _ in range(3):
        eager.step()
        graph.run(1)
        _same_state(eager, graph)
    graph.run(10)
    for _ in range(10):
        eager.step()
    _same_state(eager, graph)
    assert graph.iteration == eager.iteration == 16
    assert graph.current_time == eager.current_time
    # eager steps of the same ensemble between replays: the record reads their fields
    graph.step()
    eager.step()
    graph.run(2)
    eager.step()
    eager.step()
    _same_state(eager, graph)


@pytest.mark.parametrize("key", [48, 20])
def test_resident_step_overflow_sums_each_system_directly(td3, key):
    """A node capacity too small for the forest: the resident step makes every root a leaf
    (exact sums per system), reports it, and the next recording has room."""
    ens_mod = sys.modules["tree_d_3d_amd.ensemble"]
    sizes = [700, 40, 1300]
    ens = td3.Ensemble(gaussian_systems(td3, sizes, seed=3))
    ens.compute_fields()
    ens._resident = ens_mod._Resident(ens, len(sizes) + 1, key)
    ens.graph_capacity = len(sizes) + 1
    print(f"resident step with 4 nodes of capacity, {key}-level keys", flush=True)
    ens.step_resident()
    assert ens._resident is None and ens.graph_capacity > 8
    starts = np.concatenate(([0], np.cumsum(sizes)))
    for s in range(len(sizes)):
        sl = slice(int(starts[s]), int(starts[s + 1]))
        k = sizes[s]
        pos = ens.positions[:, sl]
        want, _ = td3.direct_field(pos, pos, np.ones(k), softening=1e-3, exclude=np.arange(k))
        got = ens._potential[sl]
        assert float((got - want).abs().max() / want.abs().max()) < 1e-13
    ens.run(3)
    assert ens._resident is not None and int(ens._resident.sizes[3]) == 0


def test_resident_step_short_key_is_reported_and_recorded_again(td3):
    """Keys of too few levels for the forest's depth: exact sums per system for that step,
    then the step is recorded again with keys of every level, and replays equal step()."""
    ens_mod = sys.modules["tree_d_3d_amd.ensemble"]
    sizes = [900, 2000, 50]
    eager = td3.Ensemble(gaussian_systems(td3, sizes, seed=8))
    ens = td3.Ensemble(gaussian_systems(td3, sizes, seed=8))
    ens.compute_fields()
    ens._resident = ens_mod._Resident(ens, 1 << 16, 3)
    print("resident step with 3-level keys", flush=True)
    ens.step_resident()
    eager.step()
    assert ens._full_key and ens._resident is None
    a, b = ens._potential, eager._potential
    assert float((a - b).abs().max() / b.abs().max()) < 1e-2  # exact sums vs the walk
    ens.positions.copy_(eager.positions)
    ens.velocities.copy_(eager.velocities)
    ens.roots.copy_(eager.roots)
    ens._current = False
    ens.run(4)
    for _ in range(4):
        eager.step()
    assert ens._resident.key_levels == 48
    _same_state(eager, ens)


def test_simulation_run_graph_equals_step(td3):
    """ElectrostaticSimulation.run(graph=True) (the step recorded as a HIP graph through a
    one-system ensemble) gives step()'s states bit for bit, across calls, and keeps the
    bodies' views; other methods refuse it."""
    a = gaussian_systems(td3, [3000], seed=21)[0]
    b = gaussian_systems(td3, [3000], seed=21)[0]
    print("simulation run(graph=True)", flush=True)
    for _ in range(5):
        a.step()
    b.run(5, graph=True)
    assert torch.equal(a.positions, b.positions) and torch.equal(a.velocities, b.velocities)
    for _ in range(3):
        a.step()
    b.run(3)
    assert torch.equal(a.positions, b.positions) and torch.equal(a.velocities, b.velocities)
    # a step() between runs, then a run again: the record reads the new state
    a.step()
    b.step()
    for _ in range(2):
        a.step()
    b.run(2)
    assert torch.equal(a.positions, b.positions) and torch.equal(a.velocities, b.velocities)
    assert b.iteration == a.iteration == 11 and b.current_time == a.current_time
    assert np.array_equal(np.asarray(a.bodies[7].pos), np.asarray(b.bodies[7].pos))
    ea, eb = a.get_total_energy(), b.get_total_energy()
    assert abs(ea - eb) <= 1e-12 * abs(ea)
    c = gaussian_systems(td3, [500], seed=1, method="direct")[0]
    with pytest.raises(ValueError):
        c.run(2, graph=True)


def test_ensemble_refuses_systems_that_resolve_to_another_method(td3):
    """A default ('auto') simulation resolves by its own size and theta (simulation.py:
    567-581): the ensemble, which runs the walk, takes it only when that is 'tree'."""
    print("ensemble method check", flush=True)
    small = gaussian_systems(td3, [300], seed=5, method="auto")  # few charges: direct
    with pytest.raises(ValueError, match="tree"):
        td3.Ensemble(small)
    walk = gaussian_systems(td3, [300], seed=5, method="auto", theta=1.2,
                            direct_threshold=10)  # theta >= 1: the walk
    assert walk[0]._resolve_method() == "tree"
    td3.Ensemble(walk)
