"""One cell step (simulation/simulator.py:1-81): field update, lifecycle, mechanical relaxation.

Order: ``oxygen.update`` (deposit -> steady-state SOR -> sample), ``lifecycle_step`` (decide ->
prefix-sum slots -> place daughters; one host sync), ``relax_contacts`` (``mechanics_substeps``
x build -> gather -> integrate, no sync). The lifecycle at step n sees the oxygen of the
configuration at the start of step n and the crowding left by the mechanics of step n-1.
"""

from __future__ import annotations

from dataclasses import dataclass

from warpbiocell_amd.cells.lifecycle import LifecycleParams, lifecycle_step
from warpbiocell_amd.cells.mechanics import ContactParams, relax_contacts


@dataclass(frozen=True)
class TimeStepping:
    dt_cells: float = 0.25
    dt_mechanics: float = 0.025

    def __post_init__(self):
        if self.dt_cells <= 0.0 or self.dt_mechanics <= 0.0:
            raise ValueError("time steps must be positive")
        if self.dt_mechanics > self.dt_cells:
            raise ValueError("dt_mechanics cannot exceed dt_cells")

    @property
    def mechanics_substeps(self) -> int:
        return max(1, int(round(self.dt_cells / self.dt_mechanics)))


@dataclass(frozen=True)
class StepReport:
    daughters: int
    field_sweeps: int = 0
    field_residual: float = 0.0
    field_converged: bool = True


def cell_step(population, grid, lifecycle: LifecycleParams, contact: ContactParams, stepping: TimeStepping,
              oxygen=None, region=None, network=None) -> StepReport:
    if network is not None:
        raise NotImplementedError("the per-cell gene network is not part of the HIP hot path")
    report = StepReport(0)
    if oxygen is not None:
        f = oxygen.update(population)
        report = StepReport(0, f.sweeps, f.residual, f.converged)
    n_daughters = lifecycle_step(population, lifecycle, stepping.dt_cells)
    relax_contacts(population, grid, contact, stepping.dt_mechanics, stepping.mechanics_substeps, region)
    return StepReport(n_daughters, report.field_sweeps, report.field_residual, report.field_converged)
