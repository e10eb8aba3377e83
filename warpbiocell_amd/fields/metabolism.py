"""Several diffusible species on one grid, sharing the per-state cell densities
(fields/metabolism.py:1-111).

Oxygen is always the first species (the lifecycle reads ``oxygen_local``); glucose, when present,
is sampled into ``glucose_local``; any other species into ``population.extra_local[name]``.
Species are solved in list order, and a species may take its production source from an earlier
one (lactate from glucose: production times S/(K_S + S) of the source field). Each solve is the
HIP solve loop with the species' production and source terms (the one-launch or colour-packed
sweeps need neither, so a producing species runs the two colour passes of the reference).
"""

from __future__ import annotations

from dataclasses import dataclass

from warpbiocell_amd.fields.densities import CellDensities
from warpbiocell_amd.fields.diffusion import SolverSettings
from warpbiocell_amd.fields.oxygen import OxygenField, OxygenParams
from warpbiocell_amd.fields.scalar_field import Boundary, GridGeometry
from warpbiocell_amd.fields.species import SpeciesField, SpeciesParams


@dataclass(frozen=True)
class MetabolicReport:
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
    def __init__(self, geometry: GridGeometry, oxygen_params: OxygenParams, species: list[SpeciesParams],
                 settings: SolverSettings = SolverSettings(), boundary=(Boundary.DIRICHLET,) * 3, device=None,
                 fixed_outside=None, oxygen_state_factors: dict | None = None):
        self.densities = CellDensities(geometry, device=device)
        self.oxygen = OxygenField(geometry, oxygen_params, settings, boundary=boundary, device=device,
                                  fixed_outside=fixed_outside, densities=self.densities,
                                  uptake_state_factors=oxygen_state_factors)
        self.device = self.oxygen.device
        self.species_fields: dict[str, SpeciesField] = {"oxygen": self.oxygen}
        self._settings, self._boundary, self._fixed_outside = settings, boundary, fixed_outside
        for params in species:
            self.add_species(params)
        self.last_report: MetabolicReport | None = None

    def add_species(self, params: SpeciesParams) -> SpeciesField:
        """Append a species after oxygen (and after any species it takes its production from); the
        order of addition is the solve order."""
        if params.name in self.species_fields:
            raise ValueError(f"duplicate species {params.name!r}")
        if params.production_source is not None and params.production_source not in self.species_fields:
            raise ValueError(f"species {params.name!r}: production source {params.production_source!r} must be an "
                             f"earlier species")
        field = SpeciesField(self.geometry, params, self._settings, boundary=self._boundary, device=self.device,
                             fixed_outside=self._fixed_outside, densities=self.densities)
        self.species_fields[params.name] = field
        return field

    @property
    def geometry(self) -> GridGeometry:
        return self.oxygen.geometry

    @property
    def names(self) -> list[str]:
        return list(self.species_fields)

    @property
    def field(self):  # the oxygen ScalarField, for code that expects an OxygenField
        return self.oxygen.field

    @property
    def density(self):  # the oxygen consuming density, for code that expects an OxygenField
        return self.oxygen.density

    def numpy(self):
        return self.oxygen.numpy()

    def __getitem__(self, name: str) -> SpeciesField:
        return self.species_fields[name]

    def update(self, population) -> MetabolicReport:
        """Deposit the per-state densities once, then combine, solve and sample every species in order."""
        self.densities.deposit(population)
        reports = {}
        for name, species in self.species_fields.items():
            species.combine()
            src = species.species.production_source
            reports[name] = species.solve_steady_state(self.species_fields[src] if src else None)
            species.sample_at_cells(population)
        self.last_report = MetabolicReport(reports)
        self.oxygen.last_report = reports["oxygen"]
        return self.last_report
