"""Gradients of cell-sampled oxygen through the steady-state solve (the implicit adjoint).

There is no reference counterpart: the reference solves the steady state (fields/diffusion.py:79-123)
but does not differentiate it. ``SteadyStateOxygen`` is a ``torch.autograd.Function`` mapping
per-cell uptake rates q_c, the diffusion coefficient D and the boundary value O_b to the oxygen
sampled at the cells, S = sample(O*(q, D, O_b)), where O* solves the same discrete problem as the
SOR path:

    sum_nb O_nb - 6 O_t - h Q_t O_t / (K + O_t) = 0   on updated nodes,   O = O_b on the faces,
    h = dx^2 / D,   Q_t = sum_c w_ct q_c / dx^3 (trilinear deposit of the rates).

The backward pass scatters dL/dS to the grid (the sampling adjoint), solves A mu = g with the
symmetric operator A = diag(6 + h Q K / (K + O*)^2) - neighbours by red-black SOR, and assembles
dL/dq_c (per-cell gather), dL/dD and dL/dO_b (global sums) in ``wbc_adjoint_solve``. Only the
converged forward field is kept: the forward iterates are never stored.

Rates enter the forward through an int64 fixed-point deposit of w * q_c / q_ref (deterministic);
the gradient uses the exact trilinear weights. Grids must have Dirichlet faces and no pinned
nodes (the symmetric operator and the face term of dL/dO_b assume both).
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import torch

from warpbiocell_amd import _lib
from warpbiocell_amd.fields.diffusion import SolverSettings, UptakeKinetics, float32_threshold, solve_steady_state
from warpbiocell_amd.fields.scalar_field import GridGeometry, ScalarField


@dataclass(frozen=True)
class AdjointSettings:
    """Red-black SOR on A mu = g. ``omega=None`` takes the optimal value of the Dirichlet Laplacian,
    2 / (1 + sin(pi / (n - 1))) for the largest axis (the diagonal shift only lowers the spectral
    radius); the stopping test is max|A mu - g| / max|g|, both scaled by the diagonal."""

    omega: float | None = None
    max_sweeps: int = 20000
    tolerance: float = 1.0e-5
    check_every: int = 50

    def __post_init__(self):
        if self.omega is not None and not 0.0 < self.omega < 2.0:
            raise ValueError("omega must lie in (0, 2)")
        if self.max_sweeps < 1 or self.check_every < 1 or self.tolerance <= 0.0:
            raise ValueError("max_sweeps, check_every and tolerance must be positive")


@dataclass(frozen=True)
class ForwardState:
    """What a backward pass needs from one forward solve: the converged field, the deposited rate
    density Q / q_ref and D. Each forward allocates its own, so a later forward on the same problem
    leaves an earlier one's state intact."""

    field: ScalarField
    rate_density: torch.Tensor
    diffusion: float


class OxygenAdjointProblem:
    """Buffers for repeated forward / backward solves on one grid and one set of cell positions."""

    def __init__(self, geometry: GridGeometry, positions: torch.Tensor, michaelis_k: float, q_ref: float,
                 settings: SolverSettings = SolverSettings(omega=1.8, tolerance=1e-6, check_every=10),
                 adjoint: AdjointSettings = AdjointSettings(), device=None):
        self.device = _lib.require_gpu(device)
        self.geometry = geometry
        self.positions = positions.to(self.device, torch.float32).contiguous()
        if self.positions.ndim != 2 or self.positions.shape[1] != 3:
            raise ValueError("positions must have shape (n, 3)")
        if michaelis_k <= 0.0 or q_ref <= 0.0:
            raise ValueError("michaelis_k and q_ref must be positive")
        self.michaelis_k = float(michaelis_k)
        self.q_ref = float(q_ref)
        self.settings = settings
        self.adjoint_settings = adjoint
        shape = geometry.shape
        self._acc = torch.zeros(shape, dtype=torch.int64, device=self.device)
        self._g = torch.zeros(shape, dtype=torch.float32, device=self.device)
        self.mu = torch.zeros(shape, dtype=torch.float32, device=self.device)
        nbytes = _lib.query("wbc_adjoint_workspace_bytes", *shape)
        self._ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        self.state: ForwardState | None = None  # of the last forward
        self.last_forward = None
        self.last_backward = None

    @property
    def field(self) -> ScalarField | None:
        """The last forward's converged field."""
        return None if self.state is None else self.state.field

    @property
    def rate_density(self) -> torch.Tensor | None:
        """The last forward's deposited rate density Q / q_ref."""
        return None if self.state is None else self.state.rate_density

    @property
    def n(self) -> int:
        return self.positions.shape[0]

    def _stream(self) -> int:
        return _lib.stream_of(self.device)

    def forward(self, rates: torch.Tensor, diffusion: float, boundary_value: float) -> torch.Tensor:
        """Deposit the rates, solve from a uniform O_b start, sample at the cells."""
        g = self.geometry
        nx, ny, nz = g.shape
        s = self._stream()
        rates = rates.detach().to(self.device, torch.float32).contiguous()
        if rates.shape != (self.n,):
            raise ValueError(f"rates must have shape ({self.n},)")
        _lib.call("wbc_fill_zero_i64", self._acc.data_ptr(), self._acc.numel(), s)
        _lib.call("wbc_deposit_rates_i64", self.positions.data_ptr(), rates.data_ptr(), self.n, 1.0 / self.q_ref,
                  g.origin[0], g.origin[1], g.origin[2], 1.0 / g.dx, nx, ny, nz, self._acc.data_ptr(), s)
        rate_density = torch.empty(g.shape, dtype=torch.float32, device=self.device)
        _lib.call("wbc_fixed_to_float", self._acc.data_ptr(), g.node_count, 1.0 / g.voxel_volume,
                  rate_density.data_ptr(), s)
        field = ScalarField.create(g, float(boundary_value), device=self.device)
        kin = UptakeKinetics(float(diffusion), self.q_ref, self.michaelis_k)
        report = solve_steady_state(field, rate_density, kin, self.settings)
        out = torch.empty(self.n, dtype=torch.float32, device=self.device)
        _lib.call("wbc_sample_trilinear", field.values.data_ptr(), nx, ny, nz, self.positions.data_ptr(), self.n,
                  g.origin[0], g.origin[1], g.origin[2], 1.0 / g.dx, out.data_ptr(), s)
        self.last_forward = report
        self.state = ForwardState(field, rate_density, float(diffusion))
        return out

    def backward(self, grad_sample: torch.Tensor, state: ForwardState | None = None
                 ) -> tuple[torch.Tensor, float, float]:
        """(dL/dq, dL/dD, dL/dO_b) for dL/dS = grad_sample, at the solution of the forward whose
        ``state`` is given (default: the last forward)."""
        state = self.state if state is None else state
        if state is None:
            raise RuntimeError("backward needs a forward solve first")
        g = self.geometry
        nx, ny, nz = g.shape
        s = self._stream()
        go = grad_sample.detach().to(self.device, torch.float32).contiguous()
        self._g.zero_()
        _lib.call("wbc_sample_adjoint", go.data_ptr(), self.positions.data_ptr(), self.n, g.origin[0], g.origin[1],
                  g.origin[2], 1.0 / g.dx, nx, ny, nz, self._g.data_ptr(), s)
        h = g.dx ** 2 / state.diffusion
        a = self.adjoint_settings
        omega = a.omega if a.omega is not None else 2.0 / (1.0 + math.sin(math.pi / (max(g.shape) - 1)))
        grad_rate = torch.empty(self.n, dtype=torch.float32, device=self.device)
        sums = (ctypes.c_double * 2)()
        sweeps, residual = ctypes.c_int32(), ctypes.c_float()
        _lib.call("wbc_adjoint_solve", omega, a.max_sweeps, float32_threshold(a.tolerance), a.check_every, h, self.q_ref,
                  self.michaelis_k, None, state.field.values.data_ptr(), state.rate_density.data_ptr(), None,
                  self._g.data_ptr(), self.mu.data_ptr(), self.positions.data_ptr(), self.n, g.origin[0], g.origin[1],
                  g.origin[2], g.dx, nx, ny, nz, grad_rate.data_ptr(), ctypes.byref(sums), ctypes.byref(sweeps),
                  ctypes.byref(residual), self._ws.data_ptr(), self._ws.numel(), s)
        grad_D = -(h / state.diffusion) * sums[0]
        grad_Ob = sums[1]
        self.last_backward = (sweeps.value, residual.value)
        return grad_rate, grad_D, grad_Ob


class SteadyStateOxygen(torch.autograd.Function):
    """S = sample(O*(rates, D, O_b)); differentiable in rates, D and O_b (0-d tensors)."""

    @staticmethod
    def forward(ctx, rates, diffusion, boundary_value, problem: OxygenAdjointProblem):
        out = problem.forward(rates, float(diffusion), float(boundary_value))
        ctx.problem = problem
        ctx.state = problem.state  # this forward's own field and rate density
        ctx.shapes = (diffusion, boundary_value)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        grad_rate, grad_D, grad_Ob = ctx.problem.backward(grad_out, ctx.state)
        d, b = ctx.shapes
        gD = torch.tensor(grad_D, dtype=d.dtype, device=d.device) if torch.is_tensor(d) else None
        gB = torch.tensor(grad_Ob, dtype=b.dtype, device=b.device) if torch.is_tensor(b) else None
        return grad_rate, gD, gB, None


def sampled_oxygen(rates: torch.Tensor, diffusion: torch.Tensor, boundary_value: torch.Tensor,
                   problem: OxygenAdjointProblem) -> torch.Tensor:
    return SteadyStateOxygen.apply(rates, diffusion, boundary_value, problem)
