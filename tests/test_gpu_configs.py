"""The benchmark configurations' own inputs against the oracle, at sizes the oracle finishes
in seconds: config 2's clustered log FMM (N = 200,000), config 3's Coulomb fmm_at (order
10, unsoftened, Plummer sources, grid targets) and config 4's ensemble (Gaussian clouds of
unit charges in a magnetic field, stepped ten times, each system against the oracle's walk
and push as simulation.py:800-835 steps it)."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def relative(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))


def config2(n, seed=7):
    """Half uniform on [-1, 1]^2, half in 16 Gaussian clusters (sigma 0.02, centres uniform
    on [-0.8, 0.8]^2), charges +-1: the benchmark's generator."""
    rng = np.random.default_rng(seed)
    half = n // 2
    centres = rng.uniform(-0.8, 0.8, (2, 16))
    per = (n - half) // 16
    clusters = (centres[:, :, None] + rng.normal(0.0, 0.02, (2, 16, per))).reshape(2, -1)
    rest = n - half - 16 * per
    pos = np.concatenate((rng.uniform(-1, 1, (2, half + rest)), clusters), axis=1).clip(-1, 1)
    return np.ascontiguousarray(pos), rng.choice([-1.0, 1.0], n)


def plummer(n, seed=3):
    """2D Plummer-like radial profile r = 0.1 sqrt(u / (1 - u)), uniform angle, kept inside
    [-1, 1]^2; charges U[-1, 1]."""
    rng = np.random.default_rng(seed)
    pts = np.empty((2, 0))
    while pts.shape[1] < n:
        u = rng.uniform(0, 1, n)
        r = 0.1 * np.sqrt(u / (1 - u))
        a = rng.uniform(0, 2 * np.pi, n)
        cand = np.stack((r * np.cos(a), r * np.sin(a)))
        pts = np.concatenate((pts, cand[:, np.all(np.abs(cand) <= 1, axis=0)]), axis=1)
    return np.ascontiguousarray(pts[:, :n]), rng.uniform(-1, 1, n)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config2_fmm_matches_oracle(td3, oracle, precision):
    pos, q = config2(200_000)
    print(f"config 2 N=200000 {precision}", flush=True)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), kernel="log", order=12,
                            precision=precision)
    phi, field = tree.fmm(theta=0.5, softening=1e-4)
    ref = oracle.OracleTree(pos, q, (0, 0, 1), kernel="log", order=12)
    rphi, rfield = ref.fmm(theta=0.5, softening=1e-4)
    assert tree.fmm_pairs == ref.fmm_pairs
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert relative(phi, rphi) < tol and relative(field, rfield) < tol


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config3_coulomb_fmm_at_matches_oracle(td3, oracle, precision):
    pos, q = plummer(200_000)
    g = np.linspace(-1, 1, 256)
    gx, gy = np.meshgrid(g, g)
    targets = np.ascontiguousarray(np.stack((gx.ravel(), gy.ravel())))
    print(f"config 3 sources=200000 targets=256^2 order 10 {precision}", flush=True)
    tree = td3.FlatQuadtree(pos, q, td3.Quad(0, 0, 1), order=10, precision=precision)
    phi, field = tree.fmm_at(targets, theta=0.5)
    ref = oracle.OracleTree(pos, q, (0, 0, 1), order=10)
    rphi, rfield = ref.fmm_at(targets, theta=0.5)
    assert tree.fmm_at_pairs == tuple(len(x) for x in ref.dual_lists)
    tol = 1e-12 if precision == "fp64" else 1e-6
    assert relative(phi, rphi) < tol and relative(field, rfield) < tol


class _OracleSystem:
    """One system stepped by the oracle: simulation.py:800-835, with the root growth of
    simulation.py:625-639 before every tree and the walk's fields at the particles."""

    def __init__(self, oracle, pos, q, root, *, theta, eps, k_e, dt, b_z):
        n = pos.shape[1]
        self.o, self.q, self.root = oracle, q, root
        self.theta, self.eps, self.k_e, self.dt, self.b_z = theta, eps, k_e, dt, b_z
        self.pos, self.vel, self.forces = pos.copy(), np.zeros((2, n)), np.zeros((2, n))
        self.masses, self.fixed, self.own = np.ones(n), np.zeros(n, dtype=bool), np.arange(n)
        self.e = None

    def _grow(self):
        xmid, ymid, half = self.root
        lo = np.minimum(self.pos.min(axis=1), [xmid - half, ymid - half])
        hi = np.maximum(self.pos.max(axis=1), [xmid + half, ymid + half])
        if np.any(lo < (xmid - half, ymid - half)) or np.any(hi > (xmid + half, ymid + half)):
            centre = lo / 2 + hi / 2
            self.root = (float(centre[0]), float(centre[1]),
                         float(np.max(hi / 2 - lo / 2) * (1 + 1e-12)))

    def _electric(self):
        self._grow()
        tree = self.o.OracleTree(self.pos, self.q, self.root, leaf_capacity=8, max_depth=48)
        _, field = tree.evaluate(self.pos, exclude=self.own, theta=self.theta,
                                 softening=self.eps)
        return self.k_e * self.q * field

    def step(self):
        if self.e is None:
            self.e = self._electric()
        args = (self.masses, self.q, self.fixed, self.dt, self.b_z)
        self.o.push(self.pos, self.vel, self.forces, self.e, *args, first_half=True)
        self.e = self._electric()
        self.o.push(self.pos, self.vel, self.forces, self.e, *args, first_half=False)

    def state(self):
        return self.pos, self.vel, self.forces


STATE = ("positions", "velocities", "forces")


def test_config4_ensemble_matches_oracle(td3, oracle):
    sizes = [1024, 3000, 11000, 40000, 65536]
    rng = np.random.default_rng(4)
    pos = rng.normal(0.0, 0.1, (2, sum(sizes))).clip(-0.999, 0.999)
    n = pos.shape[1]
    roots = np.tile([0.0, 0.0, 1.0], (len(sizes), 1))
    kw = dict(theta=0.5, softening=1e-3, k_e=1.0, dt=1e-4, B_z=1.0)
    start = np.concatenate(([0], np.cumsum(sizes)))
    slices = [slice(start[s], start[s + 1]) for s in range(len(sizes))]

    def ensemble():
        return td3.Ensemble.from_arrays(pos, np.zeros((2, n)), np.ones(n), np.ones(n), sizes,
                                        roots, **kw)

    def systems():
        return [_OracleSystem(oracle, pos[:, sl], np.ones(sl.stop - sl.start), (0.0, 0.0, 1.0),
                              theta=0.5, eps=1e-3, k_e=1.0, dt=1e-4, b_z=1.0) for sl in slices]

    steps = 10
    # every step from the oracle's state: one step's difference is the walk's rounding
    print(f"config 4 ensemble sizes={sizes}: {steps} steps, each from the oracle's state",
          flush=True)
    ens, ref = ensemble(), systems()
    for step in range(steps):
        ens.step()
        for r in ref:
            r.step()
        for name, k in zip(STATE, range(3)):
            got = getattr(ens, name).cpu().numpy()
            for s, (sl, r) in enumerate(zip(slices, ref)):
                assert relative(got[:, sl], r.state()[k]) < 1e-12, (step, s, name)
        for name, k in zip(STATE, range(3)):
            getattr(ens, name).copy_(torch.from_numpy(
                np.concatenate([r.state()[k] for r in ref], axis=1)))
        ens._current = False
    # free steps: the walk's rounding differences (1e-13 per evaluation) grow along the
    # trajectories through close encounters, about tenfold a step in the larger systems
    # (after ten steps 1e-6 in the forces at N = 40000), so three are compared
    free = 3
    print(f"config 4 free run, {free} steps", flush=True)
    ens, ref = ensemble(), systems()
    ens.run(free, graph=False)
    got = {name: getattr(ens, name).cpu().numpy() for name in STATE}
    for s, (sl, r) in enumerate(zip(slices, ref)):
        for _ in range(free):
            r.step()
        for name, w in zip(STATE, r.state()):
            assert relative(got[name][:, sl], w) < 1e-8, (s, name)
    # the recorded (HIP graph) steps give the same states as the direct ones
    rec = ensemble()
    rec.run(free, graph=True)
    for name in STATE:
        assert torch.equal(getattr(rec, name), getattr(ens, name))
