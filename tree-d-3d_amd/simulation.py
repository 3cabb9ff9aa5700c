"""Electrostatic N-body dynamics in a plane with the fields on the GPU.

The reference's ``ElectrostaticSimulation`` (simulation.py:202-940): kick-drift-kick with an
exact uniform-B_z drift, one cached field evaluation per step, energies and momenta.  The
particle state lives in GPU tensors (the reference keeps it on the host and copies it to
the device every step, simulation.py:698); the push is one hand-written kernel
(``td3_push``) and the fields come from the package's direct sum, walk or FMM.  Absorbing
walls and objects (simgeometry) and MPI sharing are not part of this build.
"""

from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from . import _abi
from ._utils import _finite, _host
from .constants import K_E
from .flat_tree import COULOMB_ORDER, FlatQuadtree, direct_field
from .quad import Quad

DIRECT_THRESHOLD = 1000
_METHODS = ("auto", "direct", "tree", "fmm")


class ChargedBody:
    """A particle or weighted macroparticle (the reference's bodies.py:20-100).

    Mass and charge are per physical particle; ``weight`` multiplies both.  A simulation
    adopts its bodies: afterwards ``pos``, ``vel``, ``force``, ``mass``, ``charge`` and
    ``fixed`` read and write the body's column of the simulation's GPU arrays.  Reading
    returns a host copy (assign to the attribute to change the state).
    """

    def __init__(self, x, y, vx=0.0, vy=0.0, mass=1.0, charge=1.0, *, weight=1.0,
                 fixed=False):
        if not np.all(np.isfinite([x, y, vx, vy])):
            raise ValueError("Particle position and velocity must be finite")
        self.weight = _finite("weight", weight, positive=True)
        total_mass = _finite("mass", mass, positive=True) * self.weight
        total_charge = _finite("charge", charge) * self.weight
        _finite("weighted mass", total_mass, positive=True)
        _finite("weighted charge", total_charge)
        if fixed and (vx or vy):
            raise ValueError("A fixed source must have zero velocity")
        self._sim = None
        self._index = 0
        self._own = {"pos": np.array([x, y], dtype=float), "vel": np.array([vx, vy], dtype=float),
                     "force": np.zeros(2), "mass": total_mass, "charge": total_charge,
                     "fixed": bool(fixed)}

    def _bind(self, sim, index: int) -> None:
        self._sim, self._index = sim, index

    def _column(self, name):
        return getattr(self._sim, _ARRAYS[name])

    def _get(self, name):
        if self._sim is None:
            return self._own[name]
        col = self._column(name)
        if col.dim() == 2:
            return _host(col[:, self._index])
        value = col[self._index].item()
        return bool(value) if name == "fixed" else float(value)

    def _set(self, name, value):
        if self._sim is None:
            self._own[name] = np.array(value, dtype=float) if name in _VECTORS else value
            return
        col = self._column(name)
        if col.dim() == 2:
            col[:, self._index] = torch.as_tensor(np.asarray(value, dtype=float), device=col.device)
        else:
            col[self._index] = value
            if name == "fixed":
                self._sim._fixed_u8[self._index] = int(bool(value))

    pos = property(lambda self: self._get("pos"), lambda self, v: self._set("pos", v))
    vel = property(lambda self: self._get("vel"), lambda self, v: self._set("vel", v))
    force = property(lambda self: self._get("force"), lambda self, v: self._set("force", v))
    mass = property(lambda self: self._get("mass"), lambda self, v: self._set("mass", v))
    charge = property(lambda self: self._get("charge"), lambda self, v: self._set("charge", v))
    fixed = property(lambda self: self._get("fixed"), lambda self, v: self._set("fixed", v))

    def reset_force(self) -> None:
        self.force = np.zeros(2)

    def add_force(self, fx: float, fy: float) -> None:
        self.force = self.force + (fx, fy)


_ARRAYS = {"pos": "positions", "vel": "velocities", "force": "forces", "mass": "masses",
           "charge": "charges", "fixed": "fixed"}
_VECTORS = ("pos", "vel", "force")


class ElectrostaticSimulation:
    """Isolated, nonrelativistic electrostatic N-body dynamics in a plane (GPU state).

    Same constructor and conventions as the reference (simulation.py:415-508): ``method``
    ``"auto"`` sums directly for ``theta == 0`` or at most ``direct_threshold`` charged
    particles, else uses the FMM for ``theta < 1`` (every kernel here runs on the GPU, the
    Coulomb FMM included), else the tree walk.
    """

    def __init__(self, bodies: Sequence[ChargedBody], bounds: Quad, theta: float = 0.5,
                 k_e: float = K_E, softening: float = 0.0, dt: float = 1e-12, B_z: float = 0.0,
                 *, leaf_capacity=None, max_depth: int = 48, method: str = "auto",
                 direct_threshold: int = DIRECT_THRESHOLD, absorbing_walls=None,
                 absorbing_objects=None, kernel: str = "coulomb", order=None,
                 reference_radius: float = 1.0, symmetric: bool = False, comm=None,
                 precision: str = "fp64"):
        if absorbing_walls or absorbing_objects:
            raise NotImplementedError("absorbing walls and objects are not part of this build")
        if comm is not None:
            raise NotImplementedError("one GPU: comm must be None")
        if method not in _METHODS:
            raise ValueError(f"method must be one of {_METHODS}")
        if method == "fmm" and not 0 < theta < 1:
            raise ValueError("method='fmm' needs 0 < theta < 1")
        bodies = list(bodies)
        self.bodies = bodies
        if len({id(b) for b in bodies}) != len(bodies):
            raise ValueError("Each particle must be a distinct object")
        self.kernel = kernel
        self.order = int(order if order is not None else (18 if kernel == "log" else COULOMB_ORDER))
        self.symmetric = bool(symmetric)
        self.reference_radius = _finite("reference_radius", reference_radius, positive=True)
        if isinstance(direct_threshold, bool) or int(direct_threshold) != direct_threshold \
                or direct_threshold < 0:
            raise ValueError("direct_threshold must be a non-negative integer")
        self.direct_threshold = int(direct_threshold)
        self.theta = _finite("theta", theta, minimum=0)
        self.k_e = _finite("k_e", k_e, minimum=0)
        self.softening = _finite("softening", softening, minimum=0)
        self.dt = _finite("dt", dt, positive=True)
        self.B_z = _finite("B_z", B_z)
        self.leaf_capacity = leaf_capacity
        self.max_depth = int(max_depth)
        self.method = method
        self.precision = precision
        self.bounds = bounds
        self.current_time = 0.0
        self.iteration = 0
        dev = _abi.device()
        f64 = dict(dtype=torch.float64, device=dev)
        n = len(bodies)
        self.positions = torch.tensor(np.array([b.pos for b in bodies]).reshape(n, 2).T.copy(), **f64)
        self.velocities = torch.tensor(np.array([b.vel for b in bodies]).reshape(n, 2).T.copy(), **f64)
        self.forces = torch.tensor(np.array([b.force for b in bodies]).reshape(n, 2).T.copy(), **f64)
        self.masses = torch.tensor([b.mass for b in bodies], **f64)
        self.charges = torch.tensor([b.charge for b in bodies], **f64)
        self.fixed = torch.tensor([b.fixed for b in bodies], dtype=torch.bool, device=dev)
        self._fixed_u8 = self.fixed.to(torch.uint8)
        for i, body in enumerate(bodies):
            body._bind(self, i)
        self._electric = torch.zeros((2, n), **f64)
        self._potential = torch.zeros(n, **f64)
        self._state = None
        self._potential_done = False
        self._fixed_fields_done = False
        self._tree = None

    @property
    def num_particles(self) -> int:
        return int(self.charges.shape[0])

    # ------------------------------------------------------------------ fields
    def _resolve_method(self) -> str:
        """simulation.py:567-581 (the Coulomb FMM runs on the GPU here)."""
        if self.method != "auto":
            return self.method
        charged = int(torch.count_nonzero(self.charges))
        if self.theta == 0 or charged <= self.direct_threshold:
            return "direct"
        return "fmm" if self.theta < 1 else "tree"

    def _expand_bounds(self) -> None:
        """simulation.py:625-639: grow the root to hold every particle."""
        p = self.positions
        if not p.shape[1]:
            return
        if not bool(torch.isfinite(p).all()):
            raise ValueError("Particle positions must be finite; reduce dt")
        b = self.bounds
        lo = np.minimum(p.min(dim=1).values.cpu().numpy(), [b.xmid - b.half, b.ymid - b.half])
        hi = np.maximum(p.max(dim=1).values.cpu().numpy(), [b.xmid + b.half, b.ymid + b.half])
        if np.any(lo < (b.xmid - b.half, b.ymid - b.half)) or \
                np.any(hi > (b.xmid + b.half, b.ymid + b.half)):
            center = lo / 2 + hi / 2
            half = np.max(hi / 2 - lo / 2) * (1 + 1e-12)
            self.bounds = Quad(float(center[0]), float(center[1]), float(half))

    def _parameters(self):
        return (self.theta, self.softening, self.k_e, self.method, self.direct_threshold,
                self.leaf_capacity, self.max_depth, self.kernel, self.order,
                self.reference_radius, self.precision)

    def _fields_current(self) -> bool:
        """simulation.py:597-611: the cached fields hold while the state is unchanged."""
        if self._state is None:
            return False
        pos, q, fixed, params = self._state
        return (params == self._parameters() and torch.equal(pos, self.positions)
                and torch.equal(q, self.charges) and torch.equal(fixed, self.fixed))

    def _line_offset(self, targets) -> torch.Tensor:
        """simulation.py:669-673: raw-potential shift for the reference radius r_0."""
        others = self.charges.sum() - self.charges[targets]
        return 2 * float(np.log(self.reference_radius)) * others

    def _evaluate_targets(self, targets: torch.Tensor) -> None:
        """simulation.py:675-702: fill the cached force (and potential) at ``targets``."""
        if self.k_e == 0 or not targets.numel():
            return
        points = self.positions[:, targets].contiguous()
        if self._tree is not None:
            phi, field = self._tree.evaluate(points, exclude=targets, theta=self.theta,
                                             softening=self.softening,
                                             potential=self._potential_done)
        else:
            phi, field = direct_field(points, self.positions, self.charges,
                                      softening=self.softening, exclude=targets,
                                      kernel=self.kernel, precision=self.precision)
        if self.kernel == "log":
            phi = phi + self._line_offset(targets)
        self._potential[targets] = self.k_e * phi
        self._electric[:, targets] = self.k_e * self.charges[None, targets] * field

    def compute_fields(self, include_fixed: bool = True, potential: bool = True):
        """simulation.py:705-782: electric force and potential on every particle, cached
        while positions, charges, fixed flags and parameters are unchanged.  Stepping
        skips fixed particles as targets (their force reads zero until asked for); they
        always act as sources.  Returns the tree used, or None for direct summation."""
        want_potential = potential or self.kernel != "log"
        if want_potential and not self._potential_done and self._state is not None:
            self._state = None
        if self._fields_current():
            if include_fixed and not self._fixed_fields_done:
                self._evaluate_targets(torch.nonzero(self.charges * self.fixed).flatten())
                self._fixed_fields_done = True
            return self._tree
        self._expand_bounds()
        self._electric = torch.zeros_like(self.positions)
        self._potential = torch.zeros_like(self.charges)
        self._tree = None
        method = self._resolve_method()
        self._potential_done = want_potential or method == "direct"
        if method in ("tree", "fmm"):
            self._tree = self.build_tree()
        charged = self.charges != 0
        state = (self.positions.clone(), self.charges.clone(), self.fixed.clone(),
                 self._parameters())
        if method == "fmm":
            if self.k_e != 0 and bool(charged.any()):
                phi, field = self._tree.fmm(theta=self.theta, softening=self.softening,
                                            potential=self._potential_done)
                if self._potential_done:
                    if self.kernel == "log":
                        phi = phi + self._line_offset(slice(None))
                    self._potential = self.k_e * phi
                self._electric = self.k_e * self.charges[None, :] * field
            self._fixed_fields_done = True
            self._state = state
            return self._tree
        mobile = charged & (include_fixed | ~self.fixed)
        self._evaluate_targets(torch.nonzero(mobile).flatten())
        self._fixed_fields_done = bool(include_fixed or not bool((charged & self.fixed).any()))
        self._state = state
        return self._tree

    def build_tree(self) -> FlatQuadtree:
        """simulation.py:658-667: a tree of the particles, the root grown to hold them."""
        self._expand_bounds()
        leaf = self.leaf_capacity
        if leaf is None:
            leaf = 32 if self._resolve_method() == "fmm" else 8
        return FlatQuadtree(self.positions, self.charges, self.bounds, leaf_capacity=leaf,
                            max_depth=self.max_depth, kernel=self.kernel, order=self.order,
                            symmetric=self.symmetric, precision=self.precision)

    # ------------------------------------------------------------------ dynamics
    def step(self, dt=None):
        """simulation.py:800-835: electric half-kick, exact cyclotron drift, half-kick."""
        self.dt = _finite("dt", self.dt if dt is None else dt, positive=True)
        self.compute_fields(include_fixed=False, potential=False)
        _push(self, first_half=True)
        tree = self.compute_fields(include_fixed=False, potential=False)
        _push(self, first_half=False)
        self.current_time += self.dt
        self.iteration += 1
        return tree

    def run(self, steps: int, *, graph: bool = True):
        """``steps`` steps of `step`.  ``graph``: the step is recorded once as a HIP graph
        and replayed (`Ensemble.run` on a one-system ensemble of this simulation, kept for
        later calls), which gives the same states bit for bit; it needs the Coulomb
        Barnes-Hut walk in FP64 (``method`` resolving to ``"tree"``).  No counterpart in the
        reference, which steps on the host (simulation.py:800-835)."""
        steps = int(steps)
        if steps < 0:
            raise ValueError("steps must be non-negative")
        if not graph:
            for _ in range(steps):
                self.step()
            return self._tree
        if self.kernel != "coulomb" or self.precision != "fp64" or \
                self._resolve_method() != "tree":
            raise ValueError("graph=True records the Coulomb Barnes-Hut walk step "
                             "(kernel 'coulomb', method 'tree', precision 'fp64')")
        from .ensemble import Ensemble

        self.dt = _finite("dt", self.dt, positive=True)
        ens = getattr(self, "_graph_ensemble", None)
        if ens is None or ens.num_particles != self.num_particles:
            ens = self._graph_ensemble = Ensemble([self])
        else:
            same = (torch.equal(ens.positions, self.positions)
                    and torch.equal(ens.velocities, self.velocities)
                    and torch.equal(ens.charges, self.charges)
                    and torch.equal(ens.fixed, self.fixed) and torch.equal(ens.masses, self.masses)
                    and ens.k_e == self.k_e)
            if not same:
                ens._graph = ens._resident = None
                ens.positions.copy_(self.positions)
                ens.velocities.copy_(self.velocities)
                ens.charges.copy_(self.charges)
                ens.masses.copy_(self.masses)
                ens.fixed.copy_(self.fixed)
                ens._fixed_u8.copy_(self.fixed)
                ens.k_e = self.k_e
                ens._current = False
            b = self.bounds
            ens.roots.copy_(torch.tensor([[b.xmid, b.ymid, b.half]], dtype=torch.float64))
            ens.dt, ens.B_z, ens.theta, ens.softening = self.dt, self.B_z, self.theta, \
                self.softening
        ens.run(steps, graph=True)
        self.positions.copy_(ens.positions)
        self.velocities.copy_(ens.velocities)
        self.forces.copy_(ens.forces)
        root = ens.roots[0].tolist()
        self.bounds = Quad(*root)
        for _ in range(steps):
            self.current_time += self.dt
        self.iteration += steps
        self._state = None  # fields are evaluated again when asked for
        self._tree = None
        return None

    # ------------------------------------------------------------------ diagnostics
    def get_kinetic_energy(self) -> float:
        return float(0.5 * torch.sum(self.masses * torch.sum(self.velocities ** 2, dim=0)))

    def get_potential_energy(self, exact: bool = False) -> float:
        """simulation.py:853-916: 0.5 sum q phi from the field evaluation, or all pairs."""
        if self.k_e == 0 or not self.num_particles:
            return 0.0
        if not exact:
            self.compute_fields()
            return float(0.5 * torch.dot(self.charges, self._potential))
        everyone = torch.arange(self.num_particles, device=self.positions.device)
        phi, _ = direct_field(self.positions, self.positions, self.charges,
                              softening=self.softening, exclude=everyone, kernel=self.kernel)
        if self.kernel == "log":
            phi = phi + 2 * float(np.log(self.reference_radius)) * (self.charges.sum() - self.charges)
        return float(0.5 * self.k_e * torch.dot(self.charges, phi))

    def get_total_energy(self, exact: bool = False) -> float:
        return self.get_kinetic_energy() + self.get_potential_energy(exact)

    def get_total_charge(self) -> float:
        return float(self.charges.sum())

    def get_center_of_mass(self) -> np.ndarray:
        total = float(self.masses.sum())
        if not total:
            return np.zeros(2)
        return _host(self.positions @ self.masses) / total

    def get_momentum(self) -> np.ndarray:
        return _host(self.velocities @ self.masses)

    def get_angular_momentum(self) -> float:
        x, y = self.positions
        vx, vy = self.velocities
        return float(torch.sum(self.masses * (x * vy - y * vx)))


def _push(sim: ElectrostaticSimulation, first_half: bool) -> None:
    """The kick(-drift) of simulation.py:812-835 as one kernel over the particles."""
    n = sim.num_particles
    if not n:
        return
    _abi.call("td3_push", 1 if first_half else 0, n, _abi.ptr(sim.positions),
              _abi.ptr(sim.velocities), _abi.ptr(sim.forces), _abi.ptr(sim._electric),
              _abi.ptr(sim.masses), _abi.ptr(sim.charges), _abi.ptr(sim._fixed_u8),
              float(sim.dt), float(sim.B_z), _abi.stream())
