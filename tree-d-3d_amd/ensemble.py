"""Many independent electrostatic systems stepped together on one GPU.

The reference steps one `ElectrostaticSimulation` at a time (simulation.py:800-835), each
with its own tree (simulation.py:641-667).  An `Ensemble` holds S systems of uneven size
in one set of device arrays (the systems one after another) and takes every step for all
of them at once:

* one forest build (`td3_forest_build`): every system's tree, identical to the tree the
  system would get alone, in one sort and one topology pass;
* one upward pass for the moments of all trees;
* one Barnes-Hut walk launch over every particle of every system, each wavefront
  starting from the roots of its targets' systems (`td3_walk` with ``target_root``);
* one push launch (`td3_push`) for every particle.

`Ensemble.run(steps, graph=True)` takes the same steps through a HIP graph: the step's
launches are recorded once (torch's `CUDAGraph` on the library's stream) and replayed, so
that a step costs one graph launch and no host work.  The recorded step reads nothing back
to the host: the forest is built at a fixed node capacity with its counts and level
starts on the device (`td3_forest_build_nosync`), the moments read their level ranges
there (`td3_coulomb_moments_nosync`), and every array is allocated before the recording.
It computes the same trees, fields and states as `step`, bit for bit.

Each system keeps the reference's semantics: its root grows to hold its particles
(simulation.py:625-639), its fields are cached between the end of one step and the
start of the next (one evaluation per step), fixed particles act as sources only, and
its energy is ``K + 1/2 sum q phi`` (simulation.py:837-916).  The Coulomb kernel and the
walk (``method="tree"``) are what an ensemble runs; the systems share ``theta``,
``softening``, ``k_e``, ``dt``, ``B_z``, ``leaf_capacity`` and ``max_depth``.
"""

from __future__ import annotations

from typing import Sequence

import numpy as np
import torch

from . import _abi
from ._utils import _check, _finite, _host
from .flat_tree import COULOMB_ORDER, FlatQuadtree, _Arrays
from .simulation import ElectrostaticSimulation


class _Forest(FlatQuadtree):
    """The Coulomb trees of all systems as one node set (node s is the root of system s)."""

    def __init__(self, positions: torch.Tensor, charges: torch.Tensor, system: torch.Tensor,
                 system_start: torch.Tensor, roots: torch.Tensor, leaf_capacity: int,
                 max_depth: int):
        self.kernel, self.order, self.symmetric, self.comm = "coulomb", COULOMB_ORDER, False, None
        self.precision = "fp64"
        self._like_torch = True
        self._host_cache: dict = {}
        self._d = _Arrays(pos=positions, charge=charges)
        self.leaf_capacity, self.max_depth = int(leaf_capacity), int(max_depth)
        self._traversals: dict = {}
        self.fmm_pairs = (0, 0)
        self._system = system.to(torch.int32).contiguous()
        self._system_start = system_start.to(torch.int64).contiguous()
        self._roots = roots.to(torch.float64).contiguous()
        self._root = (0.0, 0.0, 1.0)
        self._build()
        self._sorted_charge()
        self._moments()

    def _build(self) -> None:
        pos = self._d["pos"]
        dev, n = pos.device, pos.shape[1]
        systems = self._roots.shape[0]
        capacity = 4 * (2 * n // self.leaf_capacity + 16 * systems)
        out = np.zeros(3, dtype=np.int64)
        levels_host = np.zeros(67, dtype=np.int64)
        while True:
            f64 = dict(dtype=torch.float64, device=dev)
            i64 = dict(dtype=torch.int64, device=dev)
            center = torch.empty((capacity, 2), **f64)
            half = torch.empty(capacity, **f64)
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
            code = _abi.lib().td3_forest_build(
                _abi.ptr(pos), n, systems, _abi.ptr(self._system), _abi.ptr(self._system_start),
                _abi.ptr(self._roots), self.leaf_capacity, self.max_depth, capacity,
                _abi.ptr(center), _abi.ptr(half), _abi.ptr(start), _abi.ptr(end),
                _abi.ptr(child), _abi.ptr(depth), _abi.ptr(levels), _abi.ptr(perm),
                _abi.ptr(sorted_pos), _abi.ptr(info), _abi.ptr(ws), nbytes, out.ctypes.data,
                levels_host.ctypes.data, _abi.stream(),
            )
            if code == _abi.TD3_ERR_CAPACITY:
                capacity = max(2 * capacity, int(out[2]))
                continue
            _abi.check(code, "td3_forest_build")
            break
        count, n_levels = int(out[0]), int(out[1])
        self.node_count = count
        self.levels = levels_host[: n_levels + 1].copy()
        self._d.update(perm=perm, node_center=center[:count], node_half=half[:count],
                       node_start=start[:count], node_end=end[:count], node_child=child[:count],
                       node_depth=depth[:count])
        self._sorted_pos = sorted_pos
        self._node_info = info[:count]
        self._depth = n_levels - 1

    def walk(self, theta: float, softening: float, potential: bool = True):
        """Potential and field (raw, without k_e) at every particle of every system from
        its own system's particles, its own charge excluded: one walk launch."""
        pos, d = self._d["pos"], self._d
        n = pos.shape[1]
        phi = torch.zeros(n, dtype=torch.float64, device=pos.device)
        field = torch.zeros((2, n), dtype=torch.float64, device=pos.device)
        status = torch.zeros(1, dtype=torch.int64, device=pos.device)
        everyone = torch.arange(n, device=pos.device)
        _abi.call(
            "td3_walk", _abi.COULOMB, _abi.FP64, _abi.ptr(pos), n, _abi.ptr(d["perm"]),
            _abi.ptr(everyone), _abi.ptr(self._sorted_pos), _abi.ptr(self._scharge),
            _abi.ptr(d["perm"]), n, _abi.ptr(d["node_center"]), _abi.ptr(d["node_half"]),
            _abi.ptr(self._node_info), _abi.ptr(d["node_qabs"]), _abi.ptr(d["node_q"]),
            _abi.ptr(d["node_p"]), _abi.ptr(d["node_s"]), _abi.ptr(d["moment_re"]),
            _abi.ptr(d["moment_im"]), self.order, self._depth, _abi.ptr(self._system),
            float(theta), float(softening), 1 if potential else 0, _abi.ptr(phi),
            _abi.ptr(field), _abi.ptr(status), _abi.stream(),
        )
        _check(status)
        return phi, field

    def walk_counts(self, theta: float, softening: float) -> dict:
        """The work of `walk` (td3_walk_counts, one extra launch of the same walk): cells
        accepted and source terms, per target and per wavefront."""
        pos, d = self._d["pos"], self._d
        n = pos.shape[1]
        phi = torch.zeros(n, dtype=torch.float64, device=pos.device)
        field = torch.zeros((2, n), dtype=torch.float64, device=pos.device)
        status = torch.zeros(1, dtype=torch.int64, device=pos.device)
        counts = torch.zeros(4, dtype=torch.int64, device=pos.device)
        everyone = torch.arange(n, device=pos.device)
        _abi.call(
            "td3_walk_counts", _abi.COULOMB, _abi.FP64, _abi.ptr(pos), n, _abi.ptr(d["perm"]),
            _abi.ptr(everyone), _abi.ptr(self._sorted_pos), _abi.ptr(self._scharge),
            _abi.ptr(d["perm"]), n, _abi.ptr(d["node_center"]), _abi.ptr(d["node_half"]),
            _abi.ptr(self._node_info), _abi.ptr(d["node_qabs"]), _abi.ptr(d["node_q"]),
            _abi.ptr(d["node_p"]), _abi.ptr(d["node_s"]), _abi.ptr(d["moment_re"]),
            _abi.ptr(d["moment_im"]), self.order, self._depth, _abi.ptr(self._system),
            float(theta), float(softening), 1, _abi.ptr(phi), _abi.ptr(field),
            _abi.ptr(counts), _abi.ptr(status), _abi.stream(),
        )
        _check(status)
        c = [int(v) for v in counts.cpu()]
        return {"accepted_per_target": c[0], "terms_per_target": c[1],
                "accepted_per_wave": c[2], "sources_per_wave": c[3]}


class _Resident:
    """Every array one step of an ensemble touches, allocated once at a fixed node
    capacity, and the step itself as launches that allocate nothing and read nothing back
    (so that it can be recorded into a HIP graph)."""

    def __init__(self, ens: "Ensemble", capacity: int, key_levels: int):
        dev, n = ens.positions.device, ens.num_particles
        self.key_levels = int(key_levels)
        f64 = dict(dtype=torch.float64, device=dev)
        i64 = dict(dtype=torch.int64, device=dev)
        self.capacity = int(capacity)
        c = self.capacity
        self.center = torch.empty((c, 2), **f64)
        self.half = torch.empty(c, **f64)
        self.start = torch.empty(c, **i64)
        self.end = torch.empty(c, **i64)
        self.child = torch.empty(c, **i64)
        self.depth = torch.empty(c, **i64)
        self.info = torch.empty((c, 4), dtype=torch.int32, device=dev)
        self.levels = torch.zeros(67, **i64)
        # count, levels, need on overflow, largest need, short key, short key so far
        self.sizes = torch.zeros(6, **i64)
        self.perm = torch.empty(n, **i64)
        self.sorted_pos = torch.empty((2, n), **f64)
        self.scharge = torch.empty(n, **f64)
        self.node_q = torch.empty(c, **f64)
        self.node_qabs = torch.empty(c, **f64)
        self.node_p = torch.empty((c, 2), **f64)
        self.node_s = torch.empty((c, 3), **f64)
        self.moment = torch.zeros((1, COULOMB_ORDER + 1), **f64)
        self.ws_bytes = _abi.lib().td3_tree_workspace_bytes(n, c)
        self.ws = _abi.workspace(self.ws_bytes)
        self.phi = torch.empty(n, **f64)
        self.field = torch.empty((2, n), **f64)
        self.status = torch.zeros(1, **i64)
        self.grown = torch.zeros(1, dtype=torch.int32, device=dev)
        self.everyone = torch.arange(n, device=dev)
        # fields are written in place into these from now on
        self.kq = ens.k_e * ens.charges[None, :]
        self.electric = ens._electric.clone()
        self.potential = ens._potential.clone()

    def step(self, ens: "Ensemble") -> None:
        """simulation.py:800-835 for every system, as `Ensemble.step` with current fields."""
        n, st = ens.num_particles, _abi.stream()
        ens._push(first_half=True, electric=self.electric)
        _abi.call("td3_system_bounds", _abi.ptr(ens.positions), n, ens.systems,
                  _abi.ptr(ens.system_start), _abi.ptr(ens.roots), _abi.ptr(self.grown), st)
        _abi.call(
            "td3_forest_build_nosync", _abi.ptr(ens.positions), n, ens.systems,
            _abi.ptr(ens.system), _abi.ptr(ens.system_start), _abi.ptr(ens.roots),
            ens.leaf_capacity, ens.max_depth, self.capacity, self.key_levels,
            _abi.ptr(self.center),
            _abi.ptr(self.half), _abi.ptr(self.start), _abi.ptr(self.end), _abi.ptr(self.child),
            _abi.ptr(self.depth), _abi.ptr(self.levels), _abi.ptr(self.perm),
            _abi.ptr(self.sorted_pos), _abi.ptr(self.info), _abi.ptr(self.ws), self.ws_bytes,
            _abi.ptr(self.sizes), st)
        _abi.call("td3_permute", _abi.ptr(ens.charges), 1, n, _abi.ptr(self.perm), n,
                  _abi.ptr(self.scharge), st)
        _abi.call("td3_coulomb_moments_nosync", _abi.ptr(self.sorted_pos),
                  _abi.ptr(self.scharge), n, _abi.ptr(self.center), _abi.ptr(self.info),
                  self.capacity, _abi.ptr(self.levels), min(self.key_levels, ens.max_depth),
                  _abi.ptr(self.node_q),
                  _abi.ptr(self.node_qabs), _abi.ptr(self.node_p), _abi.ptr(self.node_s), st)
        self.phi.zero_()
        self.field.zero_()
        _abi.call(
            "td3_walk", _abi.COULOMB, _abi.FP64, _abi.ptr(ens.positions), n,
            _abi.ptr(self.perm), _abi.ptr(self.everyone), _abi.ptr(self.sorted_pos),
            _abi.ptr(self.scharge), _abi.ptr(self.perm), n, _abi.ptr(self.center),
            _abi.ptr(self.half), _abi.ptr(self.info), _abi.ptr(self.node_qabs),
            _abi.ptr(self.node_q), _abi.ptr(self.node_p), _abi.ptr(self.node_s),
            _abi.ptr(self.moment), _abi.ptr(self.moment), COULOMB_ORDER, ens.max_depth,
            _abi.ptr(ens.system), float(ens.theta), float(ens.softening), 1,
            _abi.ptr(self.phi), _abi.ptr(self.field), _abi.ptr(self.status), st)
        # the fields of compute_fields, with the same roundings, in place
        torch.mul(self.phi, ens.k_e, out=self.potential)
        torch.mul(self.kq, self.field, out=self.electric)
        self.electric.masked_fill_(ens.fixed[None, :], 0.0)
        ens._push(first_half=False, electric=self.electric)


class Ensemble:
    """S independent Coulomb systems (``ElectrostaticSimulation`` semantics with
    ``method="tree"``) advanced together; see the module docstring."""

    def __init__(self, simulations: Sequence[ElectrostaticSimulation]):
        sims = list(simulations)
        if not sims:
            raise ValueError("an ensemble needs at least one system")
        first = sims[0]
        shared = ("theta", "softening", "k_e", "dt", "B_z", "max_depth", "kernel")
        for sim in sims:
            for name in shared:
                if getattr(sim, name) != getattr(first, name):
                    raise ValueError(f"every system of an ensemble must share {name}")
            if sim.kernel != "coulomb":
                raise ValueError("an ensemble runs the Coulomb kernel")
            # an 'auto' system resolves by its own size and theta (simulation.py:567-581);
            # the ensemble takes it only when that resolution is the walk
            if sim._resolve_method() != "tree":
                raise ValueError(
                    "an ensemble runs the Barnes-Hut walk: every system's method must be "
                    f"(or resolve to) 'tree', got {sim.method!r} -> {sim._resolve_method()!r}")
        leaf = first.leaf_capacity if first.leaf_capacity is not None else 8
        if any((s.leaf_capacity if s.leaf_capacity is not None else 8) != leaf for s in sims):
            raise ValueError("every system of an ensemble must share leaf_capacity")
        dev = first.positions.device
        self._setup(
            torch.cat([s.positions for s in sims], dim=1),
            torch.cat([s.velocities for s in sims], dim=1),
            torch.cat([s.masses for s in sims]), torch.cat([s.charges for s in sims]),
            torch.cat([s.fixed for s in sims]), [s.num_particles for s in sims],
            torch.tensor([[s.bounds.xmid, s.bounds.ymid, s.bounds.half] for s in sims],
                         dtype=torch.float64, device=dev),
            theta=first.theta, softening=first.softening, k_e=first.k_e, dt=first.dt,
            B_z=first.B_z, leaf_capacity=leaf, max_depth=first.max_depth)

    @classmethod
    def from_arrays(cls, positions, velocities, masses, charges, sizes, roots, *, fixed=None,
                    theta: float = 0.5, softening: float = 0.0, k_e: float = 1.0,
                    dt: float = 1e-12, B_z: float = 0.0, leaf_capacity: int = 8,
                    max_depth: int = 48) -> "Ensemble":
        """An ensemble from the systems' arrays laid one after another: positions and
        velocities (2, N), masses and charges (N,), ``sizes`` the particle count of every
        system, ``roots`` (S, 3) their root squares (xmid, ymid, half)."""
        dev = _abi.device()
        f64 = dict(dtype=torch.float64, device=dev)
        positions = torch.as_tensor(positions, **f64)
        n = positions.shape[1]
        fixed = torch.zeros(n, dtype=torch.bool, device=dev) if fixed is None else \
            torch.as_tensor(fixed, dtype=torch.bool, device=dev)
        ens = cls.__new__(cls)
        ens._setup(positions, torch.as_tensor(velocities, **f64), torch.as_tensor(masses, **f64),
                   torch.as_tensor(charges, **f64), fixed, [int(k) for k in sizes],
                   torch.as_tensor(roots, **f64), theta=theta, softening=softening, k_e=k_e,
                   dt=dt, B_z=B_z, leaf_capacity=leaf_capacity, max_depth=max_depth)
        return ens

    def _setup(self, positions, velocities, masses, charges, fixed, sizes, roots, *, theta,
               softening, k_e, dt, B_z, leaf_capacity, max_depth) -> None:
        self.theta = _finite("theta", theta, minimum=0)
        self.softening = _finite("softening", softening, minimum=0)
        self.k_e = _finite("k_e", k_e, minimum=0)
        self.dt = _finite("dt", dt, positive=True)
        self.B_z = _finite("B_z", B_z)
        self.leaf_capacity, self.max_depth = int(leaf_capacity), int(max_depth)
        self.current_time, self.iteration = 0.0, 0
        if not sizes or min(sizes) < 1:
            raise ValueError("every system needs at least one particle")
        self.sizes = np.array(sizes, dtype=np.int64)
        self.systems = len(sizes)
        n = int(self.sizes.sum())
        if positions.shape != (2, n) or velocities.shape != (2, n) or masses.shape != (n,) \
                or charges.shape != (n,) or fixed.shape != (n,) or roots.shape != (self.systems, 3):
            raise ValueError("array shapes do not match the system sizes")
        dev = positions.device
        self.positions = positions.contiguous().clone()
        self.velocities = velocities.contiguous().clone()
        self.forces = torch.zeros_like(self.positions)
        self.masses = masses.contiguous().clone()
        self.charges = charges.contiguous().clone()
        self.fixed = fixed.contiguous().clone()
        self._fixed_u8 = self.fixed.to(torch.uint8)
        offsets = np.concatenate(([0], np.cumsum(self.sizes)))
        self.system_start = torch.tensor(offsets, dtype=torch.int64, device=dev)
        self.system = torch.repeat_interleave(
            torch.arange(self.systems, dtype=torch.int32, device=dev),
            torch.tensor(self.sizes, device=dev))
        self._index = self.system.to(torch.int64)
        self.roots = roots.contiguous().clone()
        self._electric = torch.zeros_like(self.positions)
        self._potential = torch.zeros_like(self.charges)
        self._current = False
        self.forest = None
        self._resident = None  # fixed-capacity arrays of the resident / recorded step
        self._full_key = False  # a step found the forest deeper than its key
        self._levels_seen = 0  # the deepest forest built so far (level count)
        self._graph = None
        self._graph_key = None
        self.graph_capacity = None

    @property
    def num_particles(self) -> int:
        return int(self.charges.shape[0])

    # ------------------------------------------------------------------ fields
    def _expand_roots(self) -> None:
        """simulation.py:625-639 for every system at once (one block per system): a root
        that no longer holds its particles becomes the square around them and the old
        root."""
        if not bool(torch.isfinite(self.positions).all()):
            raise ValueError("Particle positions must be finite; reduce dt")
        grown = torch.zeros(1, dtype=torch.int32, device=self.positions.device)
        _abi.call("td3_system_bounds", _abi.ptr(self.positions), self.num_particles,
                  self.systems, _abi.ptr(self.system_start), _abi.ptr(self.roots),
                  _abi.ptr(grown), _abi.stream())

    def compute_fields(self):
        """Force and potential on every particle of every system (one forest build, one
        walk), cached until the positions change.  Returns the forest."""
        if self._current:
            return self.forest
        self._expand_roots()
        self.forest = _Forest(self.positions, self.charges, self.system, self.system_start,
                              self.roots, self.leaf_capacity, self.max_depth)
        self._levels_seen = max(self._levels_seen, len(self.forest.levels) - 1)
        phi, field = self.forest.walk(self.theta, self.softening)
        mobile = ~self.fixed
        self._potential = self.k_e * phi
        self._electric = torch.where(mobile[None, :], self.k_e * self.charges[None, :] * field,
                                     torch.zeros_like(field))
        self._current = True
        return self.forest

    def autotune(self, repeats: int = 3) -> int:
        """Times the walk's launch shapes (wavefronts of 64 targets per block: 1, 2, 4, 8)
        on this ensemble's own forest and keeps the fastest for every later walk of the
        process (`td3_walk_waves`); returns it.  Every shape computes the same walk (each
        wavefront's targets and traversal are the same), so the fields do not change."""
        import time

        forest = self.compute_fields()
        best, best_t = 4, float("inf")
        for waves in (1, 2, 4, 8):
            _abi.lib().td3_walk_waves(waves)
            forest.walk(self.theta, self.softening)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(repeats):
                forest.walk(self.theta, self.softening)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            if t < best_t:
                best, best_t = waves, t
        _abi.lib().td3_walk_waves(best)
        self.walk_waves = best
        return best

    # ------------------------------------------------------------------ dynamics
    def _push(self, first_half: bool, electric: torch.Tensor | None = None) -> None:
        electric = self._electric if electric is None else electric
        _abi.call("td3_push", 1 if first_half else 0, self.num_particles,
                  _abi.ptr(self.positions), _abi.ptr(self.velocities), _abi.ptr(self.forces),
                  _abi.ptr(electric), _abi.ptr(self.masses), _abi.ptr(self.charges),
                  _abi.ptr(self._fixed_u8), float(self.dt), float(self.B_z), _abi.stream())

    def step(self, dt=None):
        """simulation.py:800-835 for every system: half-kick, exact cyclotron drift, fields
        at the new positions, half-kick."""
        self.dt = _finite("dt", self.dt if dt is None else dt, positive=True)
        self.compute_fields()
        self._push(first_half=True)
        self._current = False
        forest = self.compute_fields()
        self._push(first_half=False)
        self.current_time += self.dt
        self.iteration += 1
        return forest

    def _ready_resident(self) -> _Resident:
        """The fixed-capacity arrays of the recorded step, holding the current fields; made
        with room for twice the nodes of the current forest (or the build's own estimate,
        or the capacity an overflow asked for) and keys of four levels more than its depth
        (at least 16; every level after a too-short key)."""
        self.dt = _finite("dt", self.dt, positive=True)
        forest = self.compute_fields()
        r = self._resident
        if r is None:
            n = self.num_particles
            capacity = max(4 * (2 * n // self.leaf_capacity + 16 * self.systems),
                           2 * int(forest.node_count) if forest is not None else 0,
                           self.graph_capacity or 0)
            self.graph_capacity = capacity
            key = self.max_depth
            if self._levels_seen and not self._full_key:
                key = min(self.max_depth, max(16, self._levels_seen + 4))
            r = self._resident = _Resident(self, capacity, key)
            self._graph = None
        elif self._electric is not r.electric:
            # fields computed by step() since: the recorded step reads these arrays
            r.electric.copy_(self._electric)
            r.potential.copy_(self._potential)
        self._electric, self._potential = r.electric, r.potential
        return r

    def _after_resident(self, steps: int) -> None:
        """Bookkeeping after resident steps, then the read-back of their reports: a step
        whose nodes overflowed the capacity had its systems summed directly (each root a
        leaf: exact fields), and the arrays are made again with room for them; a
        coincident pair or a too-deep tree raises."""
        r = self._resident
        self._current, self.forest = True, None
        for _ in range(steps):  # the clock of `step`, one addition per step
            self.current_time += self.dt
        self.iteration += steps
        sizes = r.sizes.cpu()
        need, short = int(sizes[3]), int(sizes[5])
        if int(sizes[2]) == 0:
            self._levels_seen = max(self._levels_seen, int(sizes[1]))
        status = r.status.clone()
        r.status.zero_()
        if short and r.key_levels >= self.max_depth:
            raise RuntimeError("the tree needs more levels than its keys hold; lower max_depth")
        if need or short:
            if need:
                self.graph_capacity = 2 * need
            self._full_key = self._full_key or bool(short)
            self._resident, self._graph = None, None
        if not bool(torch.isfinite(self.positions).all()):
            raise ValueError("Particle positions must be finite; reduce dt")
        _check(status)

    def step_resident(self) -> None:
        """One step through the launches that `run(graph=True)` records, run directly:
        the same state as `step`, from fixed-capacity arrays with no read-back inside."""
        self._ready_resident().step(self)
        self._after_resident(1)

    def run(self, steps: int, *, graph: bool = True) -> None:
        """``steps`` steps.  ``graph``: the step is recorded once into a HIP graph (after
        one step taken directly, which loads every kernel) and replayed for the rest; the
        record is kept for later calls while ``dt``, ``B_z``, ``theta`` and ``softening``
        stay.  Reports are read back once at the end (`step_resident`).  ``graph=False``
        takes ``step()`` ``steps`` times."""
        steps = int(steps)
        if steps < 0:
            raise ValueError("steps must be non-negative")
        if not graph:
            for _ in range(steps):
                self.step()
            return
        if steps == 0:
            return
        key = (self.dt, self.B_z, self.theta, self.softening)
        if self._graph is not None and self._graph_key != key:
            self._graph = None
        done = 0
        if self._graph is None:
            self.step_resident()
            done = 1
            if steps == 1:
                return
            r = self._ready_resident()
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                r.step(self)
            self._graph, self._graph_key = g, key
        self._ready_resident()
        for _ in range(steps - done):
            self._graph.replay()
        self._after_resident(steps - done)

    # ------------------------------------------------------------------ diagnostics
    def _energies(self, potential: torch.Tensor | None) -> np.ndarray:
        """(2, S): kinetic and potential energy of every system (td3_system_energies: one
        workgroup per system, FP64, a fixed order of additions)."""
        out = torch.empty((2, self.systems), dtype=torch.float64, device=self.positions.device)
        _abi.call("td3_system_energies", _abi.ptr(self.masses), _abi.ptr(self.velocities),
                  _abi.ptr(self.charges), _abi.ptr(potential), self.num_particles, self.systems,
                  _abi.ptr(self.system_start), _abi.ptr(out), _abi.stream())
        return _host(out)

    def kinetic_energies(self) -> np.ndarray:
        return self._energies(None)[0]

    def potential_energies(self) -> np.ndarray:
        """1/2 sum q phi per system (simulation.py:853-916, from the walk's potential)."""
        self.compute_fields()
        return self._energies(self._potential)[1]

    def total_energies(self) -> np.ndarray:
        return self.kinetic_energies() + self.potential_energies()

    def system_slice(self, s: int) -> slice:
        return slice(int(self.system_start[s]), int(self.system_start[s + 1]))
