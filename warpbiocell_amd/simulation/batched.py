"""Batched replica sweep: many spheroid replicas advance one ``cell_step`` together.

The reference runs a parameter sweep (simulation/sweep.py:115-162) as one run after another, each
stepping its own population with ``cell_step`` (simulation/simulator.py:58-81). ``BatchedSweep``
takes the replicas of such a sweep, each an ordinary population, oxygen field and parameter set
with its own cell count and grid, and packs them into one set of device arrays. It then advances
all of them per call of :meth:`BatchedSweep.step`. A step is, per replica:

1. the oxygen update: deposit, steady-state solve, sample;
2. the lifecycle step: decisions, then the replica's divisions (commit_divisions, cells/lifecycle.py:
   135-179: its own slot scan, daughters appended after its live cells, inside its own capacity);
3. the contact relaxation substeps.

The solves of all replicas run under one of two schedule families (``set_schedule``, chosen by
``autotune.tune_batch``): one persistent kernel launch with grid barriers between sweeps, the whole
step captured once in a hipGraph and replayed; or wide launches, one per colour pass over every
replica's colour-packed planes, driven by the host chunk by chunk between two captured graphs
(the step's parts before and after the solve). Each replica's state after a step is
bit-identical to ``cell_step`` on that replica alone, including its solve report and daughter
count. Cell counts change on the device; :meth:`BatchedSweep.counts` reads them and
:meth:`BatchedSweep.write_back` carries them into the replicas' populations.

When a replica's daughters would exceed its capacity, none are placed for it and its step goes on
(the reference's ``commit_divisions`` raises ``CapacityError`` there instead, before the contact
substeps); the next read of the counts or reports raises ``CapacityError`` naming the replica.

Limits:

- all faces Dirichlet, no pinned nodes, no production term;
- the same number of contact substeps in every replica;
- no tissue region and no gene network.

The constructor raises ``NotImplementedError`` or ``ValueError`` for anything outside these limits.
"""

from __future__ import annotations

import ctypes
from ctypes import c_float, c_int32, c_int64, c_void_p
from dataclasses import dataclass

import torch

from warpbiocell_amd import _lib
from warpbiocell_amd.cells.lifecycle import CapacityError, LifecycleParams
from warpbiocell_amd.cells.mechanics import ContactParams
from warpbiocell_amd.cells.state import CellPopulation
from warpbiocell_amd.fields.diffusion import SolveReport, float32_threshold, solver_scalars
from warpbiocell_amd.fields.oxygen import OxygenField
from warpbiocell_amd.simulation.simulator import StepReport, TimeStepping

MAX_REPLICAS = 64
# Per-cell arrays of CellPopulation carried by the batch, in the order of wbc_batch_cells.
CELL_FIELDS = ("position", "radius", "cell_state", "age", "oxygen_local", "glucose_local", "velocity",
               "neighbor_count", "divide_flag", "division_offset", "fate_flags", "rng_state", "cell_type")


class ReplicaDesc(ctypes.Structure):
    """``wbc_replica`` of include/wbc.h."""

    _fields_ = [
        ("cell_offset", c_int32), ("capacity", c_int32), ("count", c_int32),
        ("nx", c_int32), ("ny", c_int32), ("nz", c_int32),
        ("node_offset", c_int64),
        ("origin", c_float * 3), ("inv_dx", c_float), ("inv_voxel_volume", c_float),
        ("weight", c_float * 4),
        ("omega", c_float), ("dx2_over_D", c_float), ("uptake_max", c_float), ("michaelis_k", c_float),
        ("reference", c_float), ("tolerance", c_float),
        ("max_sweeps", c_int32), ("check_every", c_int32),
        ("dt_cells", c_float), ("death_rate", c_float), ("anoxic_death_rate", c_float), ("division_rate", c_float),
        ("inhibition_threshold", c_int32),
        ("hypoxia_threshold", c_float), ("death_threshold", c_float), ("glucose_threshold", c_float),
        ("glucose_death_threshold", c_float),
        ("necrosis_requires_glucose", c_int32), ("phenotype_model", c_int32),
        ("apoptosis_rate", c_float), ("necrosis_rate", c_float),
        ("cell_size", c_float), ("query_radius", c_float), ("contact_rate", c_float), ("dt_mechanics", c_float),
        ("substeps", c_int32),
        ("placement_factor", c_float),
    ]


class BatchCells(ctypes.Structure):
    """``wbc_batch_cells`` of include/wbc.h."""

    _fields_ = [(name, c_void_p) for name in CELL_FIELDS]


@dataclass
class Replica:
    """One run of the sweep: what ``cell_step`` receives for it. ``cell_size`` is the hash-grid
    bucket edge (``NeighborGrid.cell_size``; default the contact query radius, as
    ``make_neighbor_grid``)."""

    population: CellPopulation
    oxygen: OxygenField
    lifecycle: LifecycleParams
    contact: ContactParams
    stepping: TimeStepping
    label: str = ""
    cell_size: float | None = None


def _check_replica(r: Replica) -> None:
    if r.lifecycle.uses_network:
        raise NotImplementedError("the per-cell gene network is not part of the HIP hot path")
    f = r.oxygen.field
    if any(flag != 0 for flag in f.boundary_flags):
        raise NotImplementedError("the batched solve supports Dirichlet faces only")
    if f.fixed_or_none is not None:
        raise NotImplementedError("the batched solve supports no pinned nodes")
    if r.oxygen.species.produces:
        raise NotImplementedError("the batched solve supports no production term")
    r.contact.check_substep(r.stepping.dt_mechanics)


def replica_desc(r: Replica, cell_offset: int, node_offset: int) -> ReplicaDesc:
    """The scalars ``cell_step`` would pass for this replica, rounded to float32 as ctypes does."""
    pop, ox = r.population, r.oxygen
    g = ox.geometry
    kin = ox.species.kinetics
    dx2_over_D, reference = solver_scalars(ox.field, kin)
    life = r.lifecycle
    d = ReplicaDesc()
    d.cell_offset, d.capacity, d.count = cell_offset, pop.capacity, pop.count
    d.nx, d.ny, d.nz = g.shape
    d.node_offset = node_offset
    d.origin[:] = [float(v) for v in g.origin]
    d.inv_dx, d.inv_voxel_volume = 1.0 / g.dx, 1.0 / g.voxel_volume
    d.weight[:] = [float(w) for w in ox.species.uptake_weights]
    s = ox.settings
    d.omega, d.dx2_over_D, d.uptake_max, d.michaelis_k = s.omega, dx2_over_D, kin.uptake_max, kin.michaelis_k
    d.reference, d.max_sweeps, d.check_every = reference, s.max_sweeps, s.check_every
    d.tolerance = float32_threshold(s.tolerance)
    d.dt_cells, d.death_rate = r.stepping.dt_cells, life.death_rate
    d.anoxic_death_rate, d.division_rate = life.anoxic_death_rate, life.division_rate
    d.inhibition_threshold = int(life.inhibition_threshold)
    d.hypoxia_threshold, d.death_threshold = life.hypoxia_threshold, life.death_threshold
    d.glucose_threshold, d.glucose_death_threshold = life.glucose_threshold, life.glucose_death_threshold
    d.necrosis_requires_glucose = 1 if life.necrosis_requires_glucose else 0
    d.phenotype_model = 1 if life.uses_network else 0
    d.apoptosis_rate, d.necrosis_rate = life.apoptosis_rate, life.necrosis_rate
    d.cell_size = r.cell_size if r.cell_size is not None else r.contact.query_radius
    d.query_radius, d.contact_rate = r.contact.query_radius, r.contact.rate
    d.dt_mechanics, d.substeps = r.stepping.dt_mechanics, r.stepping.mechanics_substeps
    d.placement_factor = life.placement_factor
    return d


class BatchedSweep:
    """Replica-packed device state plus the library's batch handle.

    The packed arrays are copies taken at construction. :meth:`write_back` copies the batched
    state into the replicas' own tensors, and :meth:`population_view` / :meth:`field_view` read
    it in place.
    """

    def __init__(self, replicas: list[Replica], hash_dim: int = 64, device=None):
        if not replicas:
            raise ValueError("a sweep needs at least one replica")
        if len(replicas) > MAX_REPLICAS:
            raise ValueError(f"at most {MAX_REPLICAS} replicas per batch")
        if (_lib.query("wbc_batch_struct_size", 0) != ctypes.sizeof(ReplicaDesc)
                or _lib.query("wbc_batch_struct_size", 1) != ctypes.sizeof(BatchCells)):
            raise RuntimeError("the library's batch structs do not match this binding: rebuild the library")
        self.device = _lib.require_gpu(device if device is not None else replicas[0].population.device)
        for r in replicas:
            _check_replica(r)
            if r.population.device != self.device or r.oxygen.device != self.device:
                raise RuntimeError("every replica must live on the batch's device")
        substeps = {r.stepping.mechanics_substeps for r in replicas}
        if len(substeps) != 1:
            raise NotImplementedError("every replica must use the same number of contact substeps")
        self.replicas = list(replicas)
        self.hash_dim = int(hash_dim)
        self.cell_offsets, self.node_offsets = [], []
        descs = (ReplicaDesc * len(replicas))()
        cell, node = 0, 0
        for i, r in enumerate(replicas):
            self.cell_offsets.append(cell)
            self.node_offsets.append(node)
            descs[i] = replica_desc(r, cell, node)
            cell += r.population.capacity
            node += r.oxygen.geometry.node_count
        self.slots, self.nodes = cell, node
        self._descs = descs
        with torch.cuda.device(self.device):
            self.cells = {name: torch.cat([getattr(r.population, name) for r in replicas]).contiguous()
                          for name in CELL_FIELDS}
            self.field = torch.cat([r.oxygen.values.reshape(-1) for r in replicas]).contiguous()
        self._cells_struct = BatchCells(*[self.cells[name].data_ptr() for name in CELL_FIELDS])
        nbytes = _lib.query("wbc_batch_workspace_bytes", ctypes.byref(descs), len(replicas), self.hash_dim)
        if nbytes == 0:
            raise ValueError("the replicas do not form a valid batch")
        self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        handle = c_void_p()
        _lib.call("wbc_batch_create", ctypes.byref(descs), len(replicas), self.hash_dim,
                  ctypes.byref(self._cells_struct), self.field.data_ptr(), self.workspace.data_ptr(), nbytes,
                  self.stream, ctypes.byref(handle))
        self._handle = handle
        self.schedule = (2, 0)

    @property
    def stream(self) -> int:
        return _lib.stream_of(self.device)

    def __len__(self) -> int:
        return len(self.replicas)

    def _h(self) -> c_void_p:
        if self._handle is None:
            raise RuntimeError("the batch has been closed")
        return self._handle

    def close(self) -> None:
        if getattr(self, "_handle", None) is not None:
            _lib.call("wbc_batch_destroy", self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- stepping -------------------------------------------------------------------------
    def step(self, steps: int = 1, use_graph: bool = True) -> None:
        """Run ``steps`` cell steps of every replica: enqueued without host synchronisation under a
        persistent schedule; a wide schedule waits on its solve chunks (a few behind the device)."""
        _lib.call("wbc_batch_step", self._h(), int(steps), 1 if use_graph else 0, self.stream)

    def capture(self) -> None:
        """Capture the step's hipGraph now (the first ``step(use_graph=True)`` would otherwise)."""
        _lib.call("wbc_batch_step", self._h(), 0, 1, self.stream)

    def solve(self) -> None:
        """Only the steady-state solves of every replica, from their current fields."""
        _lib.call("wbc_batch_solve", self._h(), self.stream)

    def set_schedule(self, tile: int, blocks_per_cu: int = 0) -> None:
        _lib.call("wbc_batch_set_schedule", self._h(), int(tile), int(blocks_per_cu))
        self.schedule = (int(tile), int(blocks_per_cu))

    def report(self) -> list[SolveReport]:
        """The last solve of every replica (synchronises the stream); raises if a persistent
        launch's grid barrier timed out."""
        n = len(self.replicas)
        sweeps, residual, conv = (c_int32 * n)(), (c_float * n)(), (c_int32 * n)()
        err = c_int32()
        _lib.call("wbc_batch_report", self._h(), self.stream, ctypes.byref(sweeps), ctypes.byref(residual),
                  ctypes.byref(conv), None, ctypes.byref(err))
        if err.value:
            raise RuntimeError("the persistent solve's grid barrier timed out: its results are invalid")
        return [SolveReport(sweeps[i], float(residual[i]), bool(conv[i])) for i in range(n)]

    def step_reports(self) -> list[StepReport]:
        """The last step of every replica as ``cell_step`` reports it: daughters and solve."""
        _, daughters = self.counts(with_daughters=True)
        return [StepReport(d, r.sweeps, r.residual, r.converged) for d, r in zip(daughters, self.report())]

    def counts(self, with_daughters: bool = False):
        """Every replica's current cell count (synchronises the stream), and with ``with_daughters``
        the daughters its last step placed; raises ``CapacityError`` once a replica's daughters did
        not fit its capacity."""
        n = len(self.replicas)
        counts, daughters, err = (c_int32 * n)(), (c_int32 * n)(), (c_int32 * n)()
        _lib.call("wbc_batch_cell_counts", self._h(), self.stream, ctypes.byref(counts), ctypes.byref(daughters),
                  ctypes.byref(err))
        bad = [self.replicas[i].label or str(i) for i in range(n) if err[i]]
        if bad:
            raise CapacityError(f"divisions would exceed the capacity of replica(s) {', '.join(bad)}; "
                                f"no daughters were placed for them")
        return (list(counts), list(daughters)) if with_daughters else list(counts)

    def total_sweeps(self) -> list[int]:
        n = len(self.replicas)
        tot = (c_int64 * n)()
        err = c_int32()
        _lib.call("wbc_batch_report", self._h(), self.stream, None, None, None, ctypes.byref(tot), ctypes.byref(err))
        if err.value:
            raise RuntimeError("the persistent solve's grid barrier timed out: its results are invalid")
        return list(tot)

    # ---- state ----------------------------------------------------------------------------
    def population_view(self, i: int, name: str) -> torch.Tensor:
        """Replica i's slots of the packed per-cell array ``name`` (a view)."""
        off, cap = self.cell_offsets[i], self.replicas[i].population.capacity
        return self.cells[name][off: off + cap]

    def field_view(self, i: int) -> torch.Tensor:
        g = self.replicas[i].oxygen.geometry
        off = self.node_offsets[i]
        return self.field[off: off + g.node_count].view(*g.shape)

    def write_back(self) -> None:
        """Copy the batched state into every replica's population (cell arrays and count) and
        oxygen field."""
        counts = self.counts()
        for i, r in enumerate(self.replicas):
            for name in CELL_FIELDS:
                getattr(r.population, name).copy_(self.population_view(i, name))
            r.population.count = counts[i]
            r.population.forget_deposit_order()  # an order recorded before these steps is stale
            r.oxygen.values.copy_(self.field_view(i))
