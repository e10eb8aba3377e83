"""Cell lifecycle step: death, hypoxia, contact inhibition and division with deterministic slot
allocation (cells/lifecycle.py:1-187 of the reference).

``commit_divisions`` allocates daughter slots by an inclusive prefix sum of ``divide_flag``:
the k-th dividing cell in slot order fills slot ``count + k``. It contains the step's one
host synchronisation (reading the number of daughters) and raises :class:`CapacityError`
with nothing placed when the capacity would be exceeded.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

from warpbiocell_amd import _lib


class CapacityError(RuntimeError):
    """Raised when a step would create more cells than the population was allocated for."""


@dataclass(frozen=True)
class LifecycleParams:
    division_rate: float = 0.0289
    death_rate: float = 0.0
    anoxic_death_rate: float = 0.0
    inhibition_threshold: int = 8
    hypoxia_threshold: float = 0.0
    death_threshold: float = 0.0
    glucose_threshold: float = 0.0
    glucose_death_threshold: float = 0.0
    necrosis_requires_glucose: bool = False
    placement_factor: float = 1.0
    phenotype_model: str = "rules"
    apoptosis_rate: float = 0.0
    necrosis_rate: float = 0.0

    def __post_init__(self):
        if self.division_rate < 0.0 or self.death_rate < 0.0 or self.anoxic_death_rate < 0.0:
            raise ValueError("rates must be non-negative")
        if self.apoptosis_rate < 0.0 or self.necrosis_rate < 0.0:
            raise ValueError("apoptosis_rate and necrosis_rate must be non-negative")
        if self.phenotype_model not in ("rules", "network"):
            raise ValueError("phenotype_model must be 'rules' or 'network'")
        if self.inhibition_threshold < 1:
            raise ValueError("inhibition_threshold must be at least 1")
        if self.death_threshold < 0.0 or self.hypoxia_threshold < self.death_threshold:
            raise ValueError("need 0 <= death_threshold <= hypoxia_threshold")
        if self.glucose_death_threshold < 0.0 or self.glucose_threshold < self.glucose_death_threshold:
            raise ValueError("need 0 <= glucose_death_threshold <= glucose_threshold")
        if self.placement_factor <= 0.0:
            raise ValueError("placement_factor must be positive")

    @property
    def uses_network(self) -> bool:
        return self.phenotype_model == "network"

    @classmethod
    def oxygen_dependent(cls, **overrides) -> LifecycleParams:
        values = {"hypoxia_threshold": 8.0, "death_threshold": 2.0, "anoxic_death_rate": 0.5}
        values.update(overrides)
        return cls(**values)


def decide_lifecycle(population, params: LifecycleParams, dt: float) -> None:
    """Every living cell's death / state / division-intent decision for one step (no sync)."""
    if dt <= 0.0:
        raise ValueError("dt must be positive")
    n = population.count
    if n == 0:
        return
    p = population
    _lib.call("wbc_lifecycle_decide", dt, params.death_rate, params.anoxic_death_rate, params.division_rate,
              int(params.inhibition_threshold), params.hypoxia_threshold, params.death_threshold,
              params.glucose_threshold, params.glucose_death_threshold, 1 if params.necrosis_requires_glucose else 0,
              1 if params.uses_network else 0, params.apoptosis_rate, params.necrosis_rate,
              p.cell_state.data_ptr(), p.age.data_ptr(), p.neighbor_count.data_ptr(), p.oxygen_local.data_ptr(),
              p.glucose_local.data_ptr(), p.fate_flags.data_ptr(), p.rng_state.data_ptr(), p.divide_flag.data_ptr(),
              n, p.stream)


_HIP_ERROR_OUT_OF_MEMORY = 2


def commit_divisions(population, params: LifecycleParams) -> int:
    """Append the daughters of the cells whose ``divide_flag`` is set; returns their number.

    One C call: the slot scan, the daughter count read back (the step's single host sync), the
    capacity check and the daughter placement (cells/lifecycle.py:135-179)."""
    n = population.count
    if n == 0:
        return 0
    p = population
    ws = p.scan_workspace()
    out = ctypes.c_int32(0)
    rc = _lib.load().wbc_commit_divisions(
        n, p.capacity, params.placement_factor, p.divide_flag.data_ptr(), p.division_offset.data_ptr(),
        ws.data_ptr(), ws.numel(), p.position.data_ptr(), p.radius.data_ptr(), p.cell_state.data_ptr(),
        p.cell_type.data_ptr(), p.age.data_ptr(), p.oxygen_local.data_ptr(), p.glucose_local.data_ptr(),
        p.rng_state.data_ptr(), p.velocity.data_ptr(), p.neighbor_count.data_ptr(), p.fate_flags.data_ptr(),
        ctypes.byref(out), p.stream)
    n_daughters = int(out.value)
    if rc == _HIP_ERROR_OUT_OF_MEMORY and n + n_daughters > p.capacity:
        raise CapacityError(
            f"{n_daughters} divisions would bring the population to {n + n_daughters} cells, "
            f"above the allocated capacity of {p.capacity}"
        )
    if rc != 0:
        raise RuntimeError(f"wbc_commit_divisions failed with hipError_t {rc}")
    p.count = n + n_daughters
    return n_daughters


def lifecycle_step(population, params: LifecycleParams, dt: float) -> int:
    decide_lifecycle(population, params, dt)
    return commit_divisions(population, params)
