"""Steady-state reaction-diffusion with Michaelis-Menten uptake (fields/diffusion.py:1-123).

``solve_steady_state`` runs red-black SOR sweeps in place from the field's current values
(warm start), in chunks of ``check_every`` sweeps each followed by the max-norm residual,
until the residual is below ``tolerance`` or ``max_sweeps`` is reached. The whole loop runs
inside the HIP library (``wbc_solve_steady_state``): the host checks of chunk c overlap the
sweeps of chunk c+1, which a device-side latch turns into no-ops after convergence, so the
sweeps, the field and the report equal those of the synchronous loop.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from warpbiocell_amd import _lib
from warpbiocell_amd.fields.scalar_field import ScalarField


@dataclass(frozen=True)
class UptakeKinetics:
    """``q(O) = uptake_max * O / (michaelis_k + O)`` per consuming cell; ``diffusion`` D [um^2/h]."""

    diffusion: float
    uptake_max: float
    michaelis_k: float

    def __post_init__(self):
        if self.diffusion <= 0.0:
            raise ValueError("diffusion coefficient must be positive")
        if self.uptake_max < 0.0 or self.michaelis_k <= 0.0:
            raise ValueError("uptake_max must be non-negative and michaelis_k positive")


@dataclass(frozen=True)
class SolverSettings:
    """Red-black SOR settings: relaxation ``omega`` in (0, 2), sweep budget, max-norm residual
    tolerance, sweeps between residual checks."""

    omega: float = 1.5
    max_sweeps: int = 2000
    tolerance: float = 1.0e-5
    check_every: int = 20

    def __post_init__(self):
        if not 0.0 < self.omega < 2.0:
            raise ValueError("omega must lie in (0, 2)")
        if self.max_sweeps < 1 or self.check_every < 1:
            raise ValueError("max_sweeps and check_every must be positive")
        if self.tolerance <= 0.0:
            raise ValueError("tolerance must be positive")


@dataclass(frozen=True)
class SolveReport:
    sweeps: int
    residual: float
    converged: bool


def float32_threshold(tolerance: float) -> float:
    """The smallest float32 t with t >= ``tolerance``.

    The reference reads the float32 residual back as a Python float and tests ``residual <
    tolerance`` in double (fields/diffusion.py:120-121). For a float32 r, ``r < tolerance`` holds
    exactly when ``r < t``: every float32 below t is below ``tolerance``. Rounding ``tolerance`` to
    the nearest float32 instead would round 1e-6, 1e-5 and 1e-4 down, and a residual equal to that
    rounded value would converge in the reference but not here."""
    t = np.float32(tolerance)
    if float(t) < tolerance:
        t = np.nextafter(t, np.float32(np.inf))
    return float(t)


def solver_scalars(field: ScalarField, kinetics: UptakeKinetics) -> tuple[float, float]:
    """(dx^2 / D, residual reference scale), computed in double as the reference does."""
    dx2_over_D = field.geometry.dx ** 2 / kinetics.diffusion
    reference = field.boundary_value if field.boundary_value > 0.0 else 1.0
    return dx2_over_D, reference


def solve_steady_state(field: ScalarField, density: torch.Tensor, kinetics: UptakeKinetics,
                       settings: SolverSettings, residual_buffer: torch.Tensor | None = None,
                       production: torch.Tensor | None = None, source: torch.Tensor | None = None,
                       source_k: float = 1.0) -> SolveReport:
    """Iterate SOR sweeps in place from the field's current values; see the module docstring."""
    nx, ny, nz = field.geometry.shape
    dx2_over_D, reference = solver_scalars(field, kinetics)
    bc_x, bc_y, bc_z = field.boundary_flags
    if source is not None and production is None:
        production = field.fixed_zero_like()
    ws, host = field.solver_workspace(settings.max_sweeps, settings.check_every)
    sweeps, residual, converged = ctypes.c_int32(), ctypes.c_float(), ctypes.c_int32()
    _lib.call("wbc_solve_steady_state", settings.omega, settings.max_sweeps, float32_threshold(settings.tolerance),
              settings.check_every,
              dx2_over_D, kinetics.uptake_max, kinetics.michaelis_k, reference, bc_x, bc_y, bc_z,
              _lib.ptr(field.fixed_or_none), _lib.ptr(density), _lib.ptr(production), 1 if source is not None else 0,
              _lib.ptr(source), float(source_k), field.values.data_ptr(), nx, ny, nz, ws.data_ptr(), ws.numel(),
              host.data_ptr(), ctypes.byref(sweeps), ctypes.byref(residual), ctypes.byref(converged),
              _lib.stream_of(field.device))
    if residual_buffer is not None:
        residual_buffer.fill_(residual.value)
    return SolveReport(sweeps=sweeps.value, residual=float(residual.value), converged=bool(converged.value))


def sor_sweep(field: ScalarField, density: torch.Tensor, kinetics: UptakeKinetics, omega: float, color: int,
              production: torch.Tensor | None = None, source: torch.Tensor | None = None,
              source_k: float = 1.0) -> None:
    """One colour pass (the reference's ``sor_sweep`` launch)."""
    nx, ny, nz = field.geometry.shape
    dx2_over_D, _ = solver_scalars(field, kinetics)
    bc_x, bc_y, bc_z = field.boundary_flags
    if source is not None and production is None:
        production = field.fixed_zero_like()
    _lib.call("wbc_sor_sweep", int(color), omega, dx2_over_D, kinetics.uptake_max, kinetics.michaelis_k, bc_x, bc_y,
              bc_z, _lib.ptr(field.fixed_or_none), _lib.ptr(density), _lib.ptr(production),
              1 if source is not None else 0, _lib.ptr(source), float(source_k), field.values.data_ptr(), nx, ny, nz,
              _lib.stream_of(field.device))


def steady_state_residual(field: ScalarField, density: torch.Tensor, kinetics: UptakeKinetics,
                          residual_buffer: torch.Tensor, production: torch.Tensor | None = None,
                          source: torch.Tensor | None = None, source_k: float = 1.0) -> float:
    """Zero ``residual_buffer``, launch the residual, read it back (one host sync)."""
    nx, ny, nz = field.geometry.shape
    dx2_over_D, reference = solver_scalars(field, kinetics)
    bc_x, bc_y, bc_z = field.boundary_flags
    if source is not None and production is None:
        production = field.fixed_zero_like()
    residual_buffer.zero_()
    _lib.call("wbc_steady_state_residual", dx2_over_D, kinetics.uptake_max, kinetics.michaelis_k, reference, bc_x,
              bc_y, bc_z, _lib.ptr(field.fixed_or_none), _lib.ptr(density), _lib.ptr(production),
              1 if source is not None else 0, _lib.ptr(source), float(source_k), field.values.data_ptr(), nx, ny, nz,
              residual_buffer.data_ptr(), _lib.stream_of(field.device))
    return float(residual_buffer.item())
