"""Several diffusible species on one grid that share one deposit of the per-state cell densities.

Drop-in for ``MetabolicFields`` and ``MetabolicReport`` (fields/metabolism.py:1-111 of the
reference); the external adapter contract names ``MetabolicFields.update``
(docs/integrations/opencellcomms.md:41-65), so the class, its constructor, ``add_species``,
``names``, ``[name]`` and ``update`` keep the reference's names, argument meanings and errors.

A step's update is one deposit, then for each species in its registration order: combine the
densities with the species' uptake weights, solve, sample at the cells. Oxygen is registered first
(the lifecycle reads ``oxygen_local``); glucose is sampled into ``glucose_local`` and any other
species into ``population.extra_local[name]``. A species whose production is driven by another
(lactate from glucose: production x S / (K_S + S)) must be registered after it, so that its source
field is this step's. Each solve is the HIP solve loop with the species' production and source
terms; those terms are outside the one-launch and colour-packed sweeps, so a producing species
runs the two colour passes.
"""

from __future__ import annotations

from dataclasses import dataclass

from warpbiocell_amd.fields.densities import CellDensities
from warpbiocell_amd.fields.diffusion import SolveReport, SolverSettings
from warpbiocell_amd.fields.oxygen import OxygenField, OxygenParams
from warpbiocell_amd.fields.scalar_field import Boundary, GridGeometry
from warpbiocell_amd.fields.species import SpeciesField, SpeciesParams


@dataclass(frozen=True)
class MetabolicReport:
    """The solve report of every species of one update, keyed by species name, and their totals."""

    reports: dict  # species name -> SolveReport

    @property
    def sweeps(self) -> int:
        return sum(r.sweeps for r in self.reports.values())

    @property
    def residual(self) -> float:
        return max(r.residual for r in self.reports.values())

    @property
    def converged(self) -> bool:
        return all(r.converged for r in self.reports.values())


class MetabolicFields:
    """Oxygen plus the species in ``species``, solved in that order on one grid each step."""

    def __init__(self, geometry: GridGeometry, oxygen_params: OxygenParams, species: list[SpeciesParams],
                 settings: SolverSettings = SolverSettings(), boundary=(Boundary.DIRICHLET,) * 3, device=None,
                 fixed_outside=None, oxygen_state_factors: dict | None = None):
        self.densities = CellDensities(geometry, device=device)
        self.oxygen = OxygenField(geometry, oxygen_params, settings, boundary=boundary, device=device,
                                  fixed_outside=fixed_outside, densities=self.densities,
                                  uptake_state_factors=oxygen_state_factors)
        self.device = self.oxygen.device
        # Shared by every species added later: solver settings, faces, pinned region.
        self._common = dict(boundary=boundary, device=self.device, fixed_outside=fixed_outside,
                            densities=self.densities)
        self._settings = settings
        self.species_fields: dict[str, SpeciesField] = {"oxygen": self.oxygen}  # insertion order = solve order
        self.last_report: MetabolicReport | None = None
        for params in species:
            self.add_species(params)

    def _check_new(self, params: SpeciesParams) -> None:
        if params.name in self.species_fields:
            raise ValueError(f"species {params.name!r} is already registered")
        src = params.production_source
        if src is not None and src not in self.species_fields:
            raise ValueError(f"species {params.name!r} takes its production from {src!r}, which is not registered "
                             f"before it (a source must be solved first)")

    def add_species(self, params: SpeciesParams) -> SpeciesField:
        """Register a species to be solved after those already registered; returns its field."""
        self._check_new(params)
        added = SpeciesField(self.geometry, params, self._settings, **self._common)
        self.species_fields[params.name] = added
        return added

    @property
    def geometry(self) -> GridGeometry:
        return self.oxygen.geometry

    @property
    def names(self) -> list[str]:
        return list(self.species_fields)

    @property
    def field(self):  # oxygen's ScalarField, as OxygenField.field
        return self.oxygen.field

    @property
    def density(self):  # oxygen's consuming density, as OxygenField.density
        return self.oxygen.density

    def numpy(self):
        return self.oxygen.numpy()

    def __getitem__(self, name: str) -> SpeciesField:
        return self.species_fields[name]

    def _solve_one(self, species: SpeciesField, population) -> SolveReport:
        species.combine()
        src = species.species.production_source
        report = species.solve_steady_state(None if src is None else self.species_fields[src])
        species.sample_at_cells(population)
        return report

    def update(self, population) -> MetabolicReport:
        """One deposit, then combine, solve and sample each species in registration order."""
        self.densities.deposit(population)
        report = MetabolicReport({name: self._solve_one(sf, population) for name, sf in self.species_fields.items()})
        self.last_report = report
        self.oxygen.last_report = report.reports["oxygen"]  # where OxygenField callers read it
        return report
