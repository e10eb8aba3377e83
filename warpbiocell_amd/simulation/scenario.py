"""Device objects of one synthetic workload: the arguments ``cell_step`` receives for it.

Physics and parameters are those of ``warpbiocell_amd.workloads``. The lifecycle is the workload's
own when it names one; otherwise it has no deaths, no contact inhibition and divisions only when
the workload asks for them. Contacts use rate ``CONTACT_RATE`` with the query radius set by the
largest cell. The steps are ``DT_CELLS`` with ``DT_MECHANICS`` substeps.
"""

from __future__ import annotations

from dataclasses import dataclass

from warpbiocell_amd import workloads as W
from warpbiocell_amd.cells.lifecycle import LifecycleParams
from warpbiocell_amd.cells.mechanics import ContactParams, make_neighbor_grid
from warpbiocell_amd.cells.state import CellPopulation
from warpbiocell_amd.fields.diffusion import SolverSettings
from warpbiocell_amd.fields.oxygen import OxygenField, OxygenParams
from warpbiocell_amd.fields.scalar_field import GridGeometry
from warpbiocell_amd.simulation.simulator import TimeStepping
from warpbiocell_amd.spatial.neighbors import NeighborGrid


@dataclass
class Scenario:
    workload: W.Workload
    geometry: GridGeometry
    params: OxygenParams
    settings: SolverSettings
    lifecycle: LifecycleParams
    contact: ContactParams
    stepping: TimeStepping
    population: CellPopulation
    oxygen: OxygenField
    grid: NeighborGrid

    def replica(self):
        from warpbiocell_amd.simulation.batched import Replica

        return Replica(self.population, self.oxygen, self.lifecycle, self.contact, self.stepping,
                       label=self.workload.name, cell_size=self.grid.cell_size)


def geometry_of(w: W.Workload) -> GridGeometry:
    if w.grid_nodes is not None:
        n = int(w.grid_nodes)
        dx = w.box_um / (n - 1)
        half = 0.5 * (n - 1) * dx
        return GridGeometry(origin=(-half, -half, -half), dx=dx, shape=(n, n, n))
    return GridGeometry.centered_cube(w.box_um, w.dx_um)


def build(w: W.Workload, device, hash_dim: int = 128) -> Scenario:
    geom = geometry_of(w)
    params = OxygenParams(W.D_UM2_PER_H, W.UPTAKE_MMHG_UM3_PER_H, W.MICHAELIS_K, w.boundary_mmHg)
    settings = SolverSettings(omega=w.omega, tolerance=w.tolerance, max_sweeps=w.max_sweeps,
                              check_every=w.check_every)
    if w.lifecycle is not None:
        lifecycle = LifecycleParams(**w.lifecycle)
    else:
        lifecycle = LifecycleParams(division_rate=w.division_rate, death_rate=0.0, inhibition_threshold=10 ** 6)
    contact = ContactParams(stiffness=W.CONTACT_RATE, damping=1.0, max_radius=float(w.radii.max()))
    stepping = TimeStepping(dt_cells=W.DT_CELLS, dt_mechanics=W.DT_MECHANICS)
    pop = CellPopulation.from_numpy(w.positions, w.radii, capacity=w.capacity, device=device, seed=w.seed)
    ox = OxygenField(geom, params, settings, device=device)
    grid = make_neighbor_grid(contact, device, dim=hash_dim)
    return Scenario(w, geom, params, settings, lifecycle, contact, stepping, pop, ox, grid)
