"""Python face of the CPU oracle (``libwbc_oracle.so``). TEST INFRASTRUCTURE ONLY.

Loaded by the test suite, the smoke check and the CPU-baseline leg of the benchmark, never by
the product. Each function calls the C restatement in ``wbc_oracle.c`` with the same scalar
conversions the product's host code makes (double -> float32 once, at the call), so the two
see identical inputs. ``OraclePopulation`` and ``oracle_cell_step`` compose them into the
reference's step (simulation/simulator.py:58-81) on numpy arrays.
"""

from __future__ import annotations

import ctypes
import math
import subprocess
from ctypes import POINTER, c_float, c_int32, c_int64, c_uint32, c_void_p
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBRARY = HERE / "libwbc_oracle.so"

_lib = None


def build() -> Path:
    """Compile the oracle with its Makefile (always; cheap)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "-B", "libwbc_oracle.so"], check=True)
    return LIBRARY


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            build()
        _lib = ctypes.CDLL(str(LIBRARY))
        L = _lib
        f32, i32, i64, p = c_float, c_int32, c_int64, c_void_p
        sig = {
            "orc_pcg": (c_uint32, [c_uint32]),
            "orc_rand_init": (c_uint32, [i32, i32]),
            "orc_randf_step": (f32, [p]),
            "orc_sample_sphere": (None, [p, p]),
            "orc_expf": (f32, [f32]), "orc_sinf": (f32, [f32]), "orc_cosf": (f32, [f32]), "orc_acosf": (f32, [f32]),
            "orc_init_rng": (None, [p, i64, i32]),
            "orc_deposit_trilinear": (None, [p, p, i64, f32, f32, f32, f32, i32, i32, i32, p]),
            "orc_fixed_point_to_density": (None, [p, i64, f32, p]),
            "orc_combine_densities": (None, [f32, f32, f32, f32, p, i64, p]),
            "orc_sample_trilinear": (None, [p, i32, i32, i32, p, i64, f32, f32, f32, f32, p]),
            "orc_sor_sweep_f32": (None, [i32, f32, f32, f32, f32, i32, i32, i32, p, p, p, i32, p, f32, p, i32, i32, i32]),
            "orc_residual_f32": (f32, [f32, f32, f32, f32, i32, i32, i32, p, p, p, i32, p, f32, p, i32, i32, i32]),
            "orc_solve_steady_state_f32": (i32, [f32, i32, f32, i32, f32, f32, f32, f32, i32, i32, i32, p, p, p, i32,
                                                 p, f32, p, i32, i32, i32, p, p]),
            "orc_sor_sweep_f64": (None, [i32] + [ctypes.c_double] * 4 + [i32, i32, i32, p, p, p, i32, p,
                                                                         ctypes.c_double, p, i32, i32, i32]),
            "orc_residual_f64": (ctypes.c_double, [ctypes.c_double] * 4 + [i32, i32, i32, p, p, p, i32, p,
                                                                            ctypes.c_double, p, i32, i32, i32]),
            "orc_solve_steady_state_f64": (i32, [ctypes.c_double, i32, ctypes.c_double, i32] + [ctypes.c_double] * 4 +
                                           [i32, i32, i32, p, p, p, i32, p, ctypes.c_double, p, i32, i32, i32, p, p]),
            "orc_bucket_key": (i32, [f32, f32, f32, f32, i32]),
            "orc_hashgrid_build": (None, [p, i64, f32, i32, p, p]),
            "orc_contact_velocities": (None, [p, p, i64, f32, i32, p, p, f32, f32, p, p]),
            "orc_integrate_positions": (None, [p, p, i64, f32]),
            "orc_contact_substep": (None, [p, p, i64, f32, i32, f32, f32, f32, p, p, p, p]),
            "orc_confinement_velocities": (None, [p, p, i32, i32, i32, f32, f32, f32, f32, f32, p, p, i64, p]),
            "orc_scan_inclusive_i32": (None, [p, p, i64]),
            "orc_lifecycle_decide": (None, [f32, f32, f32, f32, i32, f32, f32, f32, f32, i32, i32, f32, f32,
                                            p, p, p, p, p, p, p, p, i64]),
            "orc_place_daughters": (None, [i32, f32, p, p, p, p, p, p, p, p, p, p, p, p, p, i64]),
            "orc_adjoint_sor_f64": (None, [i32] + [ctypes.c_double] * 4 + [p, p, p, p, i32, i32, i32]),
            "orc_set_threads": (None, [i32]),
            "orc_get_threads": (i32, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
    return _lib


def set_threads(n: int) -> None:
    """Threads of the oracle's parallel loops (results do not depend on it)."""
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def _p(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data


# ---- RNG and elementary functions ---------------------------------------------------------------

def init_rng(capacity: int, seed: int) -> np.ndarray:
    out = np.zeros(capacity, dtype=np.uint32)
    lib().orc_init_rng(_p(out), capacity, seed)
    return out


def pcg(v: int) -> int:
    return int(lib().orc_pcg(v & 0xFFFFFFFF))


def randf_sequence(state: int, n: int) -> tuple[np.ndarray, int]:
    s = np.array([state], dtype=np.uint32)
    out = np.array([lib().orc_randf_step(_p(s)) for _ in range(n)], dtype=np.float32)
    return out, int(s[0])


def sample_sphere(state: int) -> tuple[np.ndarray, int]:
    s = np.array([state], dtype=np.uint32)
    out = np.zeros(3, dtype=np.float32)
    lib().orc_sample_sphere(_p(s), _p(out))
    return out, int(s[0])


def expf(x: float) -> np.float32:
    return np.float32(lib().orc_expf(x))


def sinf(x: float) -> np.float32:
    return np.float32(lib().orc_sinf(x))


def cosf(x: float) -> np.float32:
    return np.float32(lib().orc_cosf(x))


def acosf(x: float) -> np.float32:
    return np.float32(lib().orc_acosf(x))


# ---- deposit / sample -------------------------------------------------------------------------------

def deposit_accumulator(positions, states, origin, dx, shape) -> np.ndarray:
    pos = np.ascontiguousarray(positions, dtype=np.float32)
    st = np.ascontiguousarray(states, dtype=np.int32)
    acc = np.zeros((4, *shape), dtype=np.int64)
    lib().orc_deposit_trilinear(_p(pos), _p(st), len(pos), origin[0], origin[1], origin[2], 1.0 / dx, *shape, _p(acc))
    return acc


def fixed_point_to_density(acc: np.ndarray, dx: float) -> np.ndarray:
    out = np.zeros(acc.shape, dtype=np.float32)
    lib().orc_fixed_point_to_density(_p(acc), acc.size, 1.0 / dx ** 3, _p(out))
    return out


def combine(weights, density: np.ndarray) -> np.ndarray:
    nodes = density[0].size
    out = np.zeros(density.shape[1:], dtype=np.float32)
    lib().orc_combine_densities(*[float(w) for w in weights], _p(np.ascontiguousarray(density)), nodes, _p(out))
    return out


def sample(field: np.ndarray, positions, origin, dx) -> np.ndarray:
    f = np.ascontiguousarray(field, dtype=np.float32)
    pos = np.ascontiguousarray(positions, dtype=np.float32)
    out = np.zeros(len(pos), dtype=np.float32)
    lib().orc_sample_trilinear(_p(f), *f.shape, _p(pos), len(pos), origin[0], origin[1], origin[2], 1.0 / dx, _p(out))
    return out


# ---- SOR ------------------------------------------------------------------------------------------

@dataclass
class SolveArgs:
    """Scalars of one steady-state problem (fields/diffusion.py:94-99)."""

    dx: float
    diffusion: float
    uptake_max: float
    michaelis_k: float
    boundary_value: float
    boundary: tuple = (0, 0, 0)

    @property
    def dx2_over_D(self) -> float:
        return self.dx ** 2 / self.diffusion

    @property
    def reference(self) -> float:
        return self.boundary_value if self.boundary_value > 0.0 else 1.0


def _dtype_suffix(dtype):
    return ("f64", np.float64) if dtype == np.float64 else ("f32", np.float32)


def _source_args(source, source_k, dt):
    """(has_source, source array, source_k): the production factor S/(K_S + S) of a source
    species (kernels/field_kernels.py:145-150), or none."""
    if source is None:
        return 0, None, 1.0
    return 1, np.ascontiguousarray(source, dtype=dt), source_k


def sor_sweep(field: np.ndarray, density: np.ndarray, a: SolveArgs, omega: float, color: int, fixed=None,
              production=None, source=None, source_k: float = 1.0) -> None:
    suf, dt = _dtype_suffix(field.dtype)
    dens = np.ascontiguousarray(density, dtype=dt)
    prod = None if production is None else np.ascontiguousarray(production, dtype=dt)
    fx = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.int32)
    hs, src, sk = _source_args(source, source_k, dt)
    getattr(lib(), f"orc_sor_sweep_{suf}")(color, omega, a.dx2_over_D, a.uptake_max, a.michaelis_k, *a.boundary,
                                            _p(fx), _p(dens), _p(prod), hs, _p(src), sk, _p(field), *field.shape)


def residual(field: np.ndarray, density: np.ndarray, a: SolveArgs, fixed=None, production=None, source=None,
             source_k: float = 1.0) -> float:
    suf, dt = _dtype_suffix(field.dtype)
    dens = np.ascontiguousarray(density, dtype=dt)
    prod = None if production is None else np.ascontiguousarray(production, dtype=dt)
    fx = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.int32)
    hs, src, sk = _source_args(source, source_k, dt)
    return float(getattr(lib(), f"orc_residual_{suf}")(a.dx2_over_D, a.uptake_max, a.michaelis_k, a.reference,
                                                      *a.boundary, _p(fx), _p(dens), _p(prod), hs, _p(src), sk,
                                                      _p(field), *field.shape))


def solve(field: np.ndarray, density: np.ndarray, a: SolveArgs, omega: float, max_sweeps: int, tolerance: float,
          check_every: int, fixed=None, production=None, source=None, source_k: float = 1.0) -> tuple[int, float, bool]:
    """fields/diffusion.py:79-123 in place on ``field`` (float32 or float64)."""
    suf, dt = _dtype_suffix(field.dtype)
    dens = np.ascontiguousarray(density, dtype=dt)
    prod = None if production is None else np.ascontiguousarray(production, dtype=dt)
    fx = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.int32)
    hs, src, sk = _source_args(source, source_k, dt)
    res = np.zeros(1, dtype=dt)
    conv = np.zeros(1, dtype=np.int32)
    if dt == np.float32:
        # fields/diffusion.py:120-121 tests the float32 residual against the double tolerance; for
        # a float32 r that is r < t with t the smallest float32 >= tolerance.
        t = np.float32(tolerance)
        tolerance = float(t if float(t) >= tolerance else np.nextafter(t, np.float32(np.inf)))
    sweeps = getattr(lib(), f"orc_solve_steady_state_{suf}")(
        omega, max_sweeps, tolerance, check_every, a.dx2_over_D, a.uptake_max, a.michaelis_k, a.reference, *a.boundary,
        _p(fx), _p(dens), _p(prod), hs, _p(src), sk, _p(field), *field.shape, _p(res), _p(conv))
    return int(sweeps), float(res[0]), bool(conv[0])


def apply_dirichlet(values: np.ndarray, boundary, value: float) -> None:
    for axis, b in enumerate(boundary):
        if b == 0:
            idx = [slice(None)] * 3
            idx[axis] = 0
            values[tuple(idx)] = value
            idx[axis] = -1
            values[tuple(idx)] = value


# ---- hash grid / mechanics ---------------------------------------------------------------------------

def hashgrid_build(positions, cell_size: float, dim: int) -> tuple[np.ndarray, np.ndarray]:
    pos = np.ascontiguousarray(positions, dtype=np.float32)
    start = np.zeros(dim ** 3 + 1, dtype=np.int32)
    ids = np.zeros(max(len(pos), 1), dtype=np.int32)
    lib().orc_hashgrid_build(_p(pos), len(pos), cell_size, dim, _p(start), _p(ids))
    return start, ids[: len(pos)]


def contact_velocities(positions, radii, cell_size: float, dim: int, query_radius: float, rate: float):
    pos = np.ascontiguousarray(positions, dtype=np.float32)
    rad = np.ascontiguousarray(radii, dtype=np.float32)
    start, ids = hashgrid_build(pos, cell_size, dim)
    vel = np.zeros_like(pos)
    cnt = np.zeros(len(pos), dtype=np.int32)
    lib().orc_contact_velocities(_p(pos), _p(rad), len(pos), cell_size, dim, _p(start), _p(ids), query_radius, rate,
                                 _p(vel), _p(cnt))
    return vel, cnt


def node_gradient(values: np.ndarray, dx: float) -> np.ndarray:
    """geometry/region.py:18-21: numpy's second-order central differences inside, first-order
    one-sided at the faces, in float64, rounded to float32; shape (nx, ny, nz, 3)."""
    grads = np.gradient(np.asarray(values, dtype=np.float64), dx, edge_order=1)
    return np.ascontiguousarray(np.stack(grads, axis=-1).astype(np.float32))


@dataclass
class OracleRegion:
    """The device half of TissueRegion (geometry/region.py:24-34) on numpy arrays."""

    sdf: np.ndarray
    gradient: np.ndarray
    origin: tuple
    dx: float

    @classmethod
    def from_sdf(cls, sdf: np.ndarray, origin, dx: float) -> "OracleRegion":
        sdf = np.ascontiguousarray(sdf, dtype=np.float32)
        return cls(sdf, node_gradient(sdf, dx), tuple(float(o) for o in origin), float(dx))

    def args(self):
        o = self.origin
        return (_p(self.sdf), _p(self.gradient), *self.sdf.shape, o[0], o[1], o[2], 1.0 / self.dx)


def confinement_velocities(region: OracleRegion, wall_rate: float, positions, radii, velocity: np.ndarray) -> None:
    """kernels/geometry_kernels.py:59-79, in place on ``velocity`` (n, 3) float32."""
    pos = np.ascontiguousarray(positions, dtype=np.float32)
    rad = np.ascontiguousarray(radii, dtype=np.float32)
    assert velocity.dtype == np.float32 and velocity.flags.c_contiguous and velocity.shape == pos.shape
    lib().orc_confinement_velocities(*region.args(), wall_rate, _p(pos), _p(rad), len(pos), _p(velocity))


def sample_sdf(region: OracleRegion, positions) -> np.ndarray:
    """kernels/geometry_kernels.py:46-56: the same clamp and lerp order as the field sampler
    (kernels/field_kernels.py:153-182), so the oracle's trilinear sampler restates both."""
    pos = np.ascontiguousarray(positions, dtype=np.float32)
    out = np.zeros(len(pos), dtype=np.float32)
    o = region.origin
    lib().orc_sample_trilinear(_p(region.sdf), *region.sdf.shape, _p(pos), len(pos), o[0], o[1], o[2],
                               1.0 / region.dx, _p(out))
    return out


def relax_contacts(positions: np.ndarray, radii: np.ndarray, cell_size: float, dim: int, query_radius: float,
                   rate: float, dt: float, substeps: int, velocity: np.ndarray, counts: np.ndarray,
                   region: OracleRegion | None = None, wall_rate: float = 0.0) -> None:
    """In place on ``positions`` (n, 3) float32 (cells/mechanics.py:75-122): per substep build,
    contact gather, the confinement push when a region is given, x += v dt."""
    n = len(positions)
    start = np.zeros(dim ** 3 + 1, dtype=np.int32)
    ids = np.zeros(max(n, 1), dtype=np.int32)
    for _ in range(substeps):
        if region is None:
            lib().orc_contact_substep(_p(positions), _p(radii), n, cell_size, dim, query_radius, rate, dt,
                                      _p(velocity), _p(counts), _p(start), _p(ids))
            continue
        if n == 0:
            return
        lib().orc_hashgrid_build(_p(positions), n, cell_size, dim, _p(start), _p(ids))
        lib().orc_contact_velocities(_p(positions), _p(radii), n, cell_size, dim, _p(start), _p(ids), query_radius,
                                     rate, _p(velocity), _p(counts))
        lib().orc_confinement_velocities(*region.args(), wall_rate, _p(positions), _p(radii), n, _p(velocity))
        lib().orc_integrate_positions(_p(positions), _p(velocity), n, dt)


def scan_inclusive(flags) -> np.ndarray:
    f = np.ascontiguousarray(flags, dtype=np.int32)
    out = np.zeros_like(f)
    lib().orc_scan_inclusive_i32(_p(f), _p(out), len(f))
    return out


# ---- population and step ----------------------------------------------------------------------------

@dataclass
class OraclePopulation:
    """numpy SoA with the layout of the reference's CellPopulation (cells/state.py:31-111)."""

    capacity: int
    count: int
    seed: int
    position: np.ndarray
    radius: np.ndarray
    cell_state: np.ndarray
    cell_type: np.ndarray
    age: np.ndarray
    oxygen_local: np.ndarray
    glucose_local: np.ndarray
    rng_state: np.ndarray
    velocity: np.ndarray
    neighbor_count: np.ndarray
    divide_flag: np.ndarray
    division_offset: np.ndarray
    fate_flags: np.ndarray
    extra: dict = field(default_factory=dict)

    @classmethod
    def from_numpy(cls, positions, radii, capacity=None, seed=0, states=None) -> OraclePopulation:
        positions = np.ascontiguousarray(positions, dtype=np.float32)
        count = len(positions)
        capacity = count if capacity is None else capacity
        if states is None:
            states = np.zeros(count, dtype=np.int32)

        def padded(v, dtype, fill=0):
            buf = np.full((capacity,) + np.asarray(v).shape[1:], fill, dtype=dtype)
            buf[:count] = v
            return buf

        return cls(capacity=capacity, count=count, seed=seed, position=padded(positions, np.float32),
                   radius=padded(np.asarray(radii, dtype=np.float32), np.float32),
                   cell_state=padded(np.asarray(states, dtype=np.int32), np.int32, fill=3),
                   cell_type=np.zeros(capacity, dtype=np.int32), age=np.zeros(capacity, dtype=np.float32),
                   oxygen_local=np.zeros(capacity, dtype=np.float32), glucose_local=np.zeros(capacity, dtype=np.float32),
                   rng_state=init_rng(capacity, seed), velocity=np.zeros((capacity, 3), dtype=np.float32),
                   neighbor_count=np.zeros(capacity, dtype=np.int32), divide_flag=np.zeros(capacity, dtype=np.int32),
                   division_offset=np.zeros(capacity, dtype=np.int32), fate_flags=np.zeros(capacity, dtype=np.int32))


def lifecycle_decide(pop: OraclePopulation, params, dt: float) -> None:
    n = pop.count
    lib().orc_lifecycle_decide(dt, params.death_rate, params.anoxic_death_rate, params.division_rate,
                               int(params.inhibition_threshold), params.hypoxia_threshold, params.death_threshold,
                               params.glucose_threshold, params.glucose_death_threshold,
                               1 if params.necrosis_requires_glucose else 0,
                               1 if params.phenotype_model == "network" else 0, params.apoptosis_rate,
                               params.necrosis_rate, _p(pop.cell_state), _p(pop.age), _p(pop.neighbor_count),
                               _p(pop.oxygen_local), _p(pop.glucose_local), _p(pop.fate_flags), _p(pop.rng_state),
                               _p(pop.divide_flag), n)


def commit_divisions(pop: OraclePopulation, params) -> int:
    n = pop.count
    if n == 0:
        return 0
    lib().orc_scan_inclusive_i32(_p(pop.divide_flag), _p(pop.division_offset), n)
    nd = int(pop.division_offset[n - 1])
    if nd == 0:
        return 0
    if n + nd > pop.capacity:
        raise RuntimeError("capacity exceeded")
    lib().orc_place_daughters(n, params.placement_factor, _p(pop.divide_flag), _p(pop.division_offset),
                              _p(pop.position), _p(pop.radius), _p(pop.cell_state), _p(pop.cell_type), _p(pop.age),
                              _p(pop.oxygen_local), _p(pop.glucose_local), _p(pop.rng_state), _p(pop.velocity),
                              _p(pop.neighbor_count), _p(pop.fate_flags), n)
    pop.count = n + nd
    return nd


@dataclass
class OracleOxygen:
    """Oxygen field state for the oracle step (fields/species.py:190-195 with oxygen weights)."""

    shape: tuple
    origin: tuple
    args: SolveArgs
    omega: float
    max_sweeps: int
    tolerance: float
    check_every: int
    values: np.ndarray = None
    density: np.ndarray = None
    weights: tuple = (1.0, 1.0, 1.0, 0.0)

    def __post_init__(self):
        if self.values is None:
            self.values = np.full(self.shape, self.args.boundary_value, dtype=np.float32)
            apply_dirichlet(self.values, self.args.boundary, self.args.boundary_value)

    def update(self, pop: OraclePopulation) -> tuple[int, float, bool]:
        acc = deposit_accumulator(pop.position[: pop.count], pop.cell_state[: pop.count], self.origin, self.args.dx,
                                  self.shape)
        self.density = combine(self.weights, fixed_point_to_density(acc, self.args.dx))
        report = solve(self.values, self.density, self.args, self.omega, self.max_sweeps, self.tolerance,
                       self.check_every)
        pop.oxygen_local[: pop.count] = sample(self.values, pop.position[: pop.count], self.origin, self.args.dx)
        return report


def oracle_cell_step(pop: OraclePopulation, lifecycle, contact, stepping, dim: int = 128, oxygen=None):
    """simulation/simulator.py:58-81 on the oracle."""
    report = (0, 0.0, True)
    if oxygen is not None:
        report = oxygen.update(pop)
    lifecycle_decide(pop, lifecycle, stepping.dt_cells)
    daughters = commit_divisions(pop, lifecycle)
    n = pop.count
    pos = np.ascontiguousarray(pop.position[:n])
    vel = np.zeros((n, 3), dtype=np.float32)
    cnt = np.zeros(n, dtype=np.int32)
    relax_contacts(pos, pop.radius[:n].copy(), contact.query_radius, dim, contact.query_radius, contact.rate,
                   stepping.dt_mechanics, stepping.mechanics_substeps, vel, cnt)
    pop.position[:n] = pos
    pop.velocity[:n] = vel
    pop.neighbor_count[:n] = cnt
    return daughters, report


# ---- FP64 adjoint oracle of the steady-state solve -------------------------------------------------
# The discrete problem of the SOR path with per-cell rates, restated in float64 with exact
# trilinear weights: F_t(O) = sum_nb O_nb - 6 O_t - h Q_t O_t/(K + O_t) on interior nodes, O = O_b on
# the faces, Q_t = sum_c w_ct q_c / dx^3. Its adjoint A mu = g, A = diag(6 + h Q K/(K+O)^2) - nb, is
# solved by orc_adjoint_sor_f64; the gradients are assembled here.

def _deposit_weights(positions, origin, dx, shape):
    """(cell index, flat node index, weight) of the deposit rule (cells outside the support dropped)."""
    u = (np.asarray(positions, dtype=np.float64) - np.asarray(origin)) / dx
    base = np.floor(u).astype(np.int64)
    inside = np.all(base >= 0, axis=1) & (base[:, 0] + 1 < shape[0]) & (base[:, 1] + 1 < shape[1]) & \
        (base[:, 2] + 1 < shape[2])
    return _corner_weights(u, base, inside, shape)


def _sample_weights(positions, origin, dx, shape):
    """The clamped sampling rule of sample_trilinear."""
    u = (np.asarray(positions, dtype=np.float64) - np.asarray(origin)) / dx
    u = np.clip(u, 0.0, np.array(shape, dtype=np.float64) - 1.0)
    base = np.minimum(np.floor(u).astype(np.int64), np.array(shape) - 2)
    return _corner_weights(u, base, np.ones(len(u), dtype=bool), shape)


def _corner_weights(u, base, inside, shape):
    frac = u - base
    cells, nodes, weights = [], [], []
    idx = np.nonzero(inside)[0]
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                w = (frac[idx, 0] if di else 1 - frac[idx, 0]) * (frac[idx, 1] if dj else 1 - frac[idx, 1]) * \
                    (frac[idx, 2] if dk else 1 - frac[idx, 2])
                flat = ((base[idx, 0] + di) * shape[1] + base[idx, 1] + dj) * shape[2] + base[idx, 2] + dk
                cells.append(idx)
                nodes.append(flat)
                weights.append(w)
    return np.concatenate(cells), np.concatenate(nodes), np.concatenate(weights)


def forward_f64(positions, rates, diffusion, boundary_value, michaelis_k, origin, dx, shape, sweeps=20000,
                tolerance=1e-12, omega=1.8, start=None):
    """Converged float64 field and the samples at the cells."""
    c, t, w = _deposit_weights(positions, origin, dx, shape)
    Q = np.zeros(int(np.prod(shape)))
    np.add.at(Q, t, w * np.asarray(rates, dtype=np.float64)[c])
    Q = (Q / dx ** 3).reshape(shape)
    field = np.full(shape, float(boundary_value)) if start is None else start.astype(np.float64).copy()
    apply_dirichlet(field, (0, 0, 0), float(boundary_value))
    a = SolveArgs(dx=dx, diffusion=diffusion, uptake_max=1.0, michaelis_k=michaelis_k, boundary_value=boundary_value)
    report = solve(field, Q, a, omega, sweeps, tolerance, 10)
    sc, st, sw = _sample_weights(positions, origin, dx, shape)
    S = np.zeros(len(positions))
    np.add.at(S, sc, sw * field.reshape(-1)[st])
    return field, Q, S, report


def optimal_omega(shape) -> float:
    """2 / (1 + sin(pi / (n - 1))) for the longest axis: the SOR relaxation of the Dirichlet
    Laplacian with the smallest spectral radius. Any omega in (0, 2) has the same fixed point, so the
    float64 reference solves use this one to converge in fewer sweeps."""
    return 2.0 / (1.0 + math.sin(math.pi / (max(shape) - 1)))


def fd_rate_gradients(positions, rates, targets, diffusion, boundary_value, michaelis_k, origin, dx, shape, cells,
                      base_field, step=1e-3, sweeps=200, omega=None) -> np.ndarray:
    """Central differences of L = sum_c (S_c - target_c)^2 with respect to the rates of ``cells``,
    step ``step`` relative: each perturbed float64 solve starts from ``base_field`` (the converged
    float64 solution) and runs the same number of sweeps, so what remains of the start's residual
    is the same in the + and - solves and cancels in their difference (to first order). The loss
    difference is summed term by term, sum (S+ - S-)(S+ + S- - 2 t): at 100k cells a full L is
    ~4e7 while one rate moves it by ~1e-7, below the rounding of L itself."""
    q0 = np.asarray(rates, dtype=np.float64)
    t = np.asarray(targets, dtype=np.float64)
    w = optimal_omega(shape) if omega is None else omega
    out = np.zeros(len(cells))
    for n, c in enumerate(cells):
        eps = step * q0[c]
        S = []
        for sign in (1.0, -1.0):
            q = q0.copy()
            q[c] += sign * eps
            S.append(forward_f64(positions, q, diffusion, boundary_value, michaelis_k, origin, dx, shape, sweeps=sweeps,
                                 tolerance=0.0, omega=w, start=base_field)[2])
        out[n] = float(np.sum((S[0] - S[1]) * (S[0] + S[1] - 2.0 * t))) / (2.0 * eps)
    return out


def adjoint_f64(positions, grad_S, field, Q, diffusion, michaelis_k, origin, dx, shape, sweeps=4000, omega=None):
    """(dL/dq_c, dL/dD, dL/dO_b) at the float64 solution ``field`` for dL/dS = grad_S; ``sweeps``
    red-black SOR sweeps on the adjoint system at ``omega`` (default optimal_omega)."""
    omega = optimal_omega(shape) if omega is None else omega
    sc, st, sw = _sample_weights(positions, origin, dx, shape)
    g = np.zeros(int(np.prod(shape)))
    np.add.at(g, st, sw * np.asarray(grad_S, dtype=np.float64)[sc])
    g = g.reshape(shape)
    h = dx ** 2 / diffusion
    mu = np.zeros(shape)
    lib().orc_adjoint_sor_f64(sweeps, omega, h, 1.0, michaelis_k, _p(np.ascontiguousarray(Q)),
                              _p(np.ascontiguousarray(field)), _p(np.ascontiguousarray(g)), _p(mu), *shape)
    inner = (slice(1, -1),) * 3
    ratio = field / (michaelis_k + field)
    dc, dt, dw = _deposit_weights(positions, origin, dx, shape)
    grad_q = np.zeros(len(positions))
    np.add.at(grad_q, dc, dw * (mu * ratio).reshape(-1)[dt])
    grad_q *= -h / dx ** 3
    grad_D = -(h / diffusion) * float(np.sum(mu[inner] * (-Q[inner] * ratio[inner])))
    face = np.ones(shape, dtype=bool)
    face[inner] = False
    nd = np.zeros(shape)
    nd[1, :, :] += 1
    nd[-2, :, :] += 1
    nd[:, 1, :] += 1
    nd[:, -2, :] += 1
    nd[:, :, 1] += 1
    nd[:, :, -2] += 1
    grad_Ob = float(g[face].sum() + np.sum(mu[inner] * nd[inner]))
    return grad_q, grad_D, grad_Ob, mu


def adjoint_check(positions, rates, targets, diffusion, boundary_value, michaelis_k, origin, dx, shape, grad_q,
                  grad_D, grad_Ob, fd_cells=64, fd_sweeps=150, seed=0) -> dict:
    """Checks of a gradient of L = sum_c (S_c - t_c)^2 (per-cell rates, D, O_b) computed elsewhere:

    - against the float64 adjoint at the converged float64 solution: the whole rate vector
      (relative 2-norm error, largest error over the largest gradient), each rate whose gradient is
      at least 1e-3 of the largest (relative errors: median, 99th percentile, maximum), D and O_b;
    - against central differences (step 1e-3 relative, fd_rate_gradients) on ``fd_cells`` rates
      drawn at random among those checked above: each parameter's error relative to its own float64
      gradient, for the float64 adjoint (which the differences pin) and for the gradient under test.
      Cells in a depleted core (oxygen near 0) have gradients down to ~1e-7 of the largest; a
      relative error of those says nothing, so they are counted but not drawn.
    """
    pos = np.asarray(positions, dtype=np.float32)
    q = np.asarray(rates, dtype=np.float64)
    t = np.asarray(targets, dtype=np.float64)
    w = optimal_omega(shape)
    field, Q, S, rep = forward_f64(pos, q, diffusion, boundary_value, michaelis_k, origin, dx, shape,
                                   tolerance=1e-13, omega=w)
    sweeps = max(1500, 12 * max(shape))  # 129^3 at the optimal omega: converged to ~1e-15 by 2,000
    gq64, gD64, gOb64, _ = adjoint_f64(pos, 2.0 * (S - t), field, Q, diffusion, michaelis_k, origin, dx, shape,
                                       sweeps=sweeps, omega=w)
    g = np.asarray(grad_q, dtype=np.float64)
    scale = float(np.abs(gq64).max())
    big = np.abs(gq64) >= 1e-3 * scale
    rel = np.abs(g[big] - gq64[big]) / np.abs(gq64[big])
    cells = np.random.default_rng(seed).choice(np.nonzero(big)[0], fd_cells, replace=False)
    fd = fd_rate_gradients(pos, q, t, diffusion, boundary_value, michaelis_k, origin, dx, shape, cells, field,
                           sweeps=fd_sweeps, omega=w)
    return {
        "fp64_forward_sweeps": int(rep[0]),
        "rates_rel_l2": float(np.linalg.norm(g - gq64) / np.linalg.norm(gq64)),
        "rates_max_err_over_max_grad": float(np.abs(g - gq64).max() / scale),
        "rates_rel_err_median": float(np.median(rel)), "rates_rel_err_p99": float(np.percentile(rel, 99)),
        "rates_rel_err_max": float(rel.max()), "rates_checked": int(big.sum()),
        "rates_below_1e-3_of_max": int((~big).sum()),
        "D_rel_err": float(abs(grad_D - gD64) / abs(gD64)), "O_b_rel_err": float(abs(grad_Ob - gOb64) / abs(gOb64)),
        "fd_cells": cells.tolist(),
        "fd_vs_fp64_rel_err": (np.abs(fd - gq64[cells]) / np.abs(gq64[cells])).tolist(),
        "gradient_vs_fd_rel_err": (np.abs(g[cells] - fd) / np.abs(fd)).tolist(),
        "fd_max_err_over_max_grad": float(np.abs(fd - gq64[cells]).max() / scale),
        "gradient_vs_fd_max_err_over_max_grad": float(np.abs(g[cells] - fd).max() / scale),
    }


# ---- population metrics (metrics/population.py) ---------------------------------------------------
# numpy restatements of the reference's host-side metrics, in its expressions and order, on the
# arrays of cells [0, count): positions (n, 3) float32, radii, states int32, per-cell values float32.
_DEAD = 3
_STATE_NAMES = ("proliferative", "quiescent", "hypoxic", "dead")


def metrics_state_counts(states: np.ndarray) -> list[int]:
    """state_counts (metrics/population.py:13-19): cells per state."""
    return [int(np.count_nonzero(states == s)) for s in range(4)]


def metrics_population_summary(positions, radii, states) -> dict:
    """population_summary without the oxygen part (metrics/population.py:22-45)."""
    counts = metrics_state_counts(states)
    x = positions
    if x.shape[0] == 0:
        centre_distance = np.zeros(0)
    else:
        centre_distance = np.linalg.norm(x - x.mean(axis=0), axis=1)
    n = int(x.shape[0])
    return {
        "cells": n,
        "living_cells": n - counts[_DEAD],
        "dead_cells": counts[_DEAD],
        "proliferative_cells": counts[0],
        "quiescent_cells": counts[1],
        "hypoxic_cells": counts[2],
        "spheroid_radius": float(np.percentile(centre_distance, 99)) if centre_distance.size else 0.0,
        "mean_radius": float(radii.mean()) if n else 0.0,
    }


def metrics_oxygen_summary(states, oxygen_local, field_values, interior, sweeps: int) -> dict:
    """oxygen_summary (metrics/population.py:48-64); interior = ScalarField.interior_mask()."""
    alive = states != _DEAD
    at_cells = oxygen_local[alive]
    grid = field_values[interior]
    return {
        "oxygen_cells_mean": float(at_cells.mean()) if at_cells.size else float("nan"),
        "oxygen_cells_min": float(at_cells.min()) if at_cells.size else float("nan"),
        "oxygen_grid_mean": float(grid.mean()),
        "oxygen_grid_min": float(grid.min()),
        "field_sweeps": sweeps,
    }


def metrics_radial_profile(positions, oxygen_local, glucose_local, states, n_bins=None, shell_width=None,
                           lactate_local=None, r=None) -> dict:
    """radial_profile (metrics/population.py:83-119). ``r`` replaces the centre distances when given
    (to bin exactly the distances another implementation computed)."""
    x = positions
    if r is None:
        r = np.linalg.norm(x - x.mean(axis=0), axis=1)
    o, s = oxygen_local, states
    r_max = float(r.max()) if r.size else 0.0
    if shell_width is not None:
        n_bins = max(1, int(np.ceil((r_max + 1e-6) / shell_width)))
        edges = shell_width * np.arange(n_bins + 1)
    else:
        n_bins = n_bins or 10
        edges = np.linspace(0.0, r_max + 1e-6, n_bins + 1)
    which = np.minimum(np.digitize(r, edges) - 1, n_bins - 1)
    out = {"r_outer": edges[1:], "cells": np.bincount(which, minlength=n_bins)}
    out["oxygen_mean"] = np.array([o[which == b].mean() if np.any(which == b) else np.nan for b in range(n_bins)])
    values = {"glucose": glucose_local}
    if lactate_local is not None:
        values["lactate"] = lactate_local
    for name, v in values.items():
        out[f"{name}_mean"] = np.array([v[which == b].mean() if np.any(which == b) else np.nan for b in range(n_bins)])
    for k, name in enumerate(_STATE_NAMES):
        out[f"fraction_{name}"] = np.array(
            [np.mean(s[which == b] == k) if np.any(which == b) else np.nan for b in range(n_bins)])
    return out


def metrics_shell_radii(profile: dict) -> dict:
    """shell_radii (metrics/population.py:122-135)."""

    def outermost(fraction):
        hit = np.where(np.nan_to_num(fraction) >= 0.5)[0]
        return float(profile["r_outer"][hit.max()]) if hit.size else 0.0

    dead = profile["fraction_dead"]
    return {
        "necrotic_radius": outermost(dead),
        "hypoxic_radius": outermost(profile["fraction_hypoxic"] + dead),
        "non_proliferative_radius": outermost(1.0 - profile["fraction_proliferative"]),
    }
