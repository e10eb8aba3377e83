"""Overdamped contact mechanics: parameters, stability rule and the substep driver
(cells/mechanics.py:20-122 of the reference).

Only ``rate = stiffness / damping`` [1/h] enters the dynamics; the neighbour query radius is
``2 * max_radius + margin`` and equals the hash-grid bucket edge. With a tissue region the
wall push ``-wall_rate * (sdf + r) * grad/|grad|`` of each cell whose surface pokes out of the
tissue is added to its contact velocity before the move.
"""

from __future__ import annotations

from dataclasses import dataclass

from warpbiocell_amd import _lib
from warpbiocell_amd.spatial.neighbors import NeighborGrid

# One overlapping pair contracts its overlap by (1 - 2 rate dt) per explicit substep; above
# rate*dt = 0.5 the pair overshoots (cells/mechanics.py:13-17).
MAX_RATE_DT = 0.5


@dataclass(frozen=True)
class ContactParams:
    stiffness: float
    damping: float
    max_radius: float
    margin: float = 0.0
    wall_rate: float = 10.0

    def __post_init__(self):
        if self.stiffness <= 0.0 or self.damping <= 0.0:
            raise ValueError("stiffness and damping must be positive")
        if self.max_radius <= 0.0:
            raise ValueError("max_radius must be positive")
        if self.margin < 0.0:
            raise ValueError("margin must be non-negative")
        if self.wall_rate < 0.0:
            raise ValueError("wall_rate must be non-negative")

    @property
    def rate(self) -> float:
        return self.stiffness / self.damping

    @property
    def query_radius(self) -> float:
        return 2.0 * self.max_radius + self.margin

    def check_substep(self, dt: float) -> None:
        if dt <= 0.0:
            raise ValueError("dt must be positive")
        if self.rate * dt > MAX_RATE_DT:
            raise ValueError(
                f"rate*dt = {self.rate * dt:.3g} exceeds {MAX_RATE_DT}: explicit contact relaxation "
                "would overshoot. Reduce dt_mechanics or stiffness/damping."
            )
        if self.wall_rate * dt > MAX_RATE_DT:
            raise ValueError(f"wall_rate*dt = {self.wall_rate * dt:.3g} exceeds {MAX_RATE_DT}: reduce dt_mechanics or wall_rate.")


def make_neighbor_grid(params: ContactParams, device=None, dim: int = 128) -> NeighborGrid:
    return NeighborGrid(cell_size=params.query_radius, dim=dim, device=device)


def _check_region(region, population) -> None:
    if region is not None and region.device != population.device:
        raise RuntimeError("population and TissueRegion must live on the same device")


def contact_substep(population, grid: NeighborGrid, params: ContactParams, dt: float, region=None) -> None:
    """Rebuild the grid, gather contact velocities and neighbour counts, add the tissue-wall push
    when a ``region`` (TissueRegion) is given, move the cells by v*dt (``dt = 0`` leaves positions
    unchanged and only refreshes the counts); cells/mechanics.py:75-109."""
    n = population.count
    if n == 0:
        return
    _check_region(region, population)
    grid.build(population)
    stream = _lib.stream_of(population.device)
    ws = grid.workspace
    _lib.call("wbc_contact_velocities", ws.data_ptr(), ws.numel(), grid.dim, grid.capacity, n, grid.cell_size,
              params.query_radius, params.rate, population.velocity.data_ptr(), population.neighbor_count.data_ptr(),
              stream)
    if region is not None:
        _lib.call("wbc_confinement_velocities", region.struct_ptr, params.wall_rate, population.position.data_ptr(),
                  population.radius.data_ptr(), n, population.velocity.data_ptr(), stream)
    _lib.call("wbc_integrate_positions", population.position.data_ptr(), population.velocity.data_ptr(), n, dt, stream)


def relax_contacts(population, grid: NeighborGrid, params: ContactParams, dt: float, substeps: int,
                   region=None) -> None:
    """``substeps`` explicit substeps, each build -> gather (+ wall push) -> x += v dt, enqueued in
    one call (cells/mechanics.py:112-122); the gather, the push and the move are one kernel."""
    params.check_substep(dt)
    n = population.count
    if n == 0 or substeps <= 0:
        return
    if population.device != grid.device:
        raise RuntimeError("population and NeighborGrid must live on the same device")
    _check_region(region, population)
    grid.reserve(population.capacity)
    ws = grid.workspace
    stream = _lib.stream_of(population.device)
    if region is None:
        _lib.call("wbc_relax_contacts", population.position.data_ptr(), population.radius.data_ptr(), n,
                  grid.cell_size, grid.dim, params.query_radius, params.rate, dt, int(substeps),
                  population.velocity.data_ptr(), population.neighbor_count.data_ptr(), grid.capacity, ws.data_ptr(),
                  ws.numel(), stream)
    else:
        _lib.call("wbc_relax_contacts_confined", population.position.data_ptr(), population.radius.data_ptr(), n,
                  grid.cell_size, grid.dim, params.query_radius, params.rate, dt, int(substeps), region.struct_ptr,
                  params.wall_rate, population.velocity.data_ptr(), population.neighbor_count.data_ptr(),
                  grid.capacity, ws.data_ptr(), ws.numel(), stream)
    grid._built_count = n
    # The last build's bucket order: the next deposit visits the cells in it (same bits, nearby
    # atomics).
    population.record_deposit_order(grid._point_ids_view())
