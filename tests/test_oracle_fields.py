"""Pin the oracle's field functions against closed forms, float64 brute force and dense solves
(the checks of the reference's tests/test_fields_computational.py and test_fields_numerical.py)."""

import numpy as np
import pytest

from tests.helpers import steady_state_dense, trilinear_deposit_f64
from warpbiocell_amd.cells.initialization import spherical_cluster
from warpbiocell_amd.fields.scalar_field import GridGeometry

MIXED = (0, 1, 0)
DIRICHLET = (0, 0, 0)


def _args(O, geom, D=2.0e3, q=5.0e4, K=3.0, ob=20.0, boundary=DIRICHLET):
    return O.SolveArgs(dx=geom.dx, diffusion=D, uptake_max=q, michaelis_k=K, boundary_value=ob, boundary=boundary)


def test_deposit_matches_float64_reference_and_conserves_cells(oracle):
    n = 300
    rng = np.random.default_rng(4)
    pos = spherical_cluster(n, 8.0, spacing_factor=1.0, jitter=0.3, seed=4)
    states = np.zeros(n, dtype=np.int32)
    states[rng.choice(n, 40, replace=False)] = 3
    geom = GridGeometry.centered_cube(300.0, 20.0)
    acc = oracle.deposit_accumulator(pos, states, geom.origin, geom.dx, geom.shape)
    density = oracle.combine((1, 1, 1, 0), oracle.fixed_point_to_density(acc, geom.dx))
    expected = trilinear_deposit_f64(pos, states != 3, geom.origin, geom.dx, geom.shape)
    np.testing.assert_allclose(density, expected, atol=1e-8)
    assert abs(density.sum() * geom.voxel_volume - (n - 40)) < 1e-3
    # Each cell contributes 2^24 +- 4 fixed-point units (8 roundings of +-0.5), in its own layer.
    per_layer = acc.reshape(4, -1).sum(axis=1)
    assert abs(per_layer[0] - (n - 40) * 2 ** 24) <= 4 * (n - 40)
    assert abs(per_layer[3] - 40 * 2 ** 24) <= 4 * 40
    assert per_layer[1] == per_layer[2] == 0


def test_deposit_ignores_cells_outside_the_grid(oracle):
    pos = np.array([[0.0, 0.0, 0.0], [500.0, 0.0, 0.0], [-100.0, 0.0, 0.0]], dtype=np.float32)
    geom = GridGeometry.centered_cube(200.0, 20.0)
    acc = oracle.deposit_accumulator(pos, np.zeros(3, np.int32), geom.origin, geom.dx, geom.shape)
    # [-100, 0, 0] sits exactly on the lower face: i0 = 0 is inside, so it is deposited; 500 is not.
    dens = oracle.fixed_point_to_density(acc, geom.dx)
    assert abs(dens.sum() * geom.voxel_volume - 2.0) < 1e-6


def test_sampling_is_exact_for_linear_fields(oracle):
    geom = GridGeometry.centered_cube(200.0, 25.0)
    x, y, z = np.meshgrid(*geom.node_coordinates(), indexing="ij")
    field = (3.0 + 0.1 * x - 0.05 * y + 0.02 * z).astype(np.float32)
    pos = np.random.default_rng(1).uniform(-90.0, 90.0, size=(200, 3)).astype(np.float32)
    out = oracle.sample(field, pos, geom.origin, geom.dx)
    expected = 3.0 + 0.1 * pos[:, 0] - 0.05 * pos[:, 1] + 0.02 * pos[:, 2]
    np.testing.assert_allclose(out, expected, atol=1e-4)


def test_sampling_clamps_positions_outside_the_grid(oracle):
    geom = GridGeometry.centered_cube(100.0, 20.0)
    field = np.full(geom.shape, 30.0, dtype=np.float32)
    pos = np.array([[1000.0, 0.0, 0.0], [0.0, -1000.0, 0.0]], dtype=np.float32)
    np.testing.assert_array_equal(oracle.sample(field, pos, geom.origin, geom.dx), 30.0)


def test_uniform_field_stays_uniform_without_consumption(oracle):
    geom = GridGeometry.centered_cube(200.0, 20.0)
    field = np.full(geom.shape, 40.0, dtype=np.float32)
    sweeps, res, conv = oracle.solve(field, np.zeros(geom.shape, np.float32), _args(oracle, geom, ob=40.0), 1.5,
                                     20000, 1e-6, 10)
    assert conv and sweeps == 10
    np.testing.assert_array_equal(field, 40.0)


def test_zero_consumption_converges_to_boundary_value_from_zero(oracle):
    geom = GridGeometry.centered_cube(300.0, 20.0)
    field = np.zeros(geom.shape, dtype=np.float32)
    oracle.apply_dirichlet(field, DIRICHLET, 38.0)
    sweeps, res, conv = oracle.solve(field, np.zeros(geom.shape, np.float32), _args(oracle, geom, ob=38.0), 1.8,
                                     20000, 1e-6, 10)
    assert conv
    assert np.abs(field - 38.0).max() < 1e-4


@pytest.mark.parametrize("omega", [1.0, 1.5, 1.9])
def test_one_dimensional_linear_uptake_matches_cosh_profile(oracle, omega):
    """Dirichlet in x, zero flux in y and z, linear uptake: O(x) = O_b cosh(kx) / cosh(kL)."""
    L, nx = 200.0, 81
    geom = GridGeometry(origin=(-L, 0.0, 0.0), dx=2 * L / (nx - 1), shape=(nx, 3, 3))
    D, kappa, K = 1.0e4, 3.0 / L, 1.0e9
    a = oracle.SolveArgs(dx=geom.dx, diffusion=D, uptake_max=kappa ** 2 * D * K, michaelis_k=K, boundary_value=1.0,
                         boundary=(0, 1, 1))
    field = np.ones(geom.shape, dtype=np.float32)
    _, _, conv = oracle.solve(field, np.ones(geom.shape, np.float32), a, omega, 20000, 1e-5, 10)
    x = geom.node_coordinates()[0]
    exact = np.cosh(kappa * x) / np.cosh(kappa * L)
    assert conv
    np.testing.assert_allclose(field[:, 1, 1], exact, atol=2e-3)
    np.testing.assert_allclose(field[:, 0, 0], field[:, 1, 1], atol=1e-6)


def test_steady_state_matches_exact_discrete_solution(oracle):
    geom = GridGeometry(origin=(0.0, 0.0, 0.0), dx=10.0, shape=(8, 7, 6))
    a = _args(oracle, geom, boundary=MIXED)
    density = np.random.default_rng(3).uniform(0.0, 2.0e-3, size=geom.shape).astype(np.float32)
    field = np.full(geom.shape, 20.0, dtype=np.float32)
    _, _, conv = oracle.solve(field, density, a, 1.5, 20000, 1e-5, 10)
    exact = steady_state_dense(np.full(geom.shape, 20.0), density, a.diffusion, a.uptake_max, a.michaelis_k, a.dx, MIXED)
    assert conv
    np.testing.assert_allclose(field, exact, atol=2e-3)
    assert exact.min() > 0.0 and exact[1:-1, :, 1:-1].max() < 20.0


def test_symmetric_consumption_gives_symmetric_field(oracle):
    geom = GridGeometry.centered_cube(400.0, 20.0)
    x, y, z = np.meshgrid(*geom.node_coordinates(), indexing="ij")
    density = np.where(x ** 2 + y ** 2 + z ** 2 < 120.0 ** 2, 2.4e-4, 0.0).astype(np.float32)
    a = _args(oracle, geom, D=7.2e6, q=1.4e8, K=3.4, ob=38.0)
    field = np.full(geom.shape, 38.0, dtype=np.float32)
    _, _, conv = oracle.solve(field, density, a, 1.8, 2000, 1e-5, 10)
    assert conv
    for axis in range(3):
        np.testing.assert_allclose(field, np.flip(field, axis=axis), atol=1e-4)
    np.testing.assert_allclose(field, np.transpose(field, (1, 2, 0)), atol=1e-4)
    assert field[10, 10, 10] < 0.75 * 38.0


def test_strong_uptake_keeps_field_non_negative(oracle):
    geom = GridGeometry.centered_cube(300.0, 20.0)
    a = _args(oracle, geom, D=7.2e6, q=1.0e12, K=3.4, ob=38.0)
    field = np.full(geom.shape, 38.0, dtype=np.float32)
    _, _, conv = oracle.solve(field, np.full(geom.shape, 1e-3, np.float32), a, 1.9, 2000, 1e-5, 10)
    assert conv and field.min() >= 0.0 and field[1:-1, 1:-1, 1:-1].max() < 1.0


def test_warm_start_needs_one_check_interval(oracle):
    geom = GridGeometry.centered_cube(400.0, 20.0)
    x, y, z = np.meshgrid(*geom.node_coordinates(), indexing="ij")
    density = np.where(x ** 2 + y ** 2 + z ** 2 < 150.0 ** 2, 2.4e-4, 0.0).astype(np.float32)
    a = _args(oracle, geom, D=7.2e6, q=1.4e8, K=3.4, ob=38.0)
    field = np.full(geom.shape, 38.0, dtype=np.float32)
    first = oracle.solve(field, density, a, 1.8, 2000, 1e-5, 10)
    second = oracle.solve(field, density, a, 1.8, 2000, 1e-5, 10)
    assert first[2] and second[2] and first[0] > 20 and second[0] == 10


def test_float32_and_float64_solves_agree(oracle):
    geom = GridGeometry.centered_cube(400.0, 20.0)
    x, y, z = np.meshgrid(*geom.node_coordinates(), indexing="ij")
    density = np.where(x ** 2 + y ** 2 + z ** 2 < 150.0 ** 2, 2.4e-4, 0.0)
    a = _args(oracle, geom, D=7.2e6, q=1.4e8, K=3.4, ob=38.0)
    f32 = np.full(geom.shape, 38.0, dtype=np.float32)
    f64 = np.full(geom.shape, 38.0, dtype=np.float64)
    oracle.solve(f32, density.astype(np.float32), a, 1.8, 600, 1e-30, 10)
    oracle.solve(f64, density, a, 1.8, 600, 1e-30, 10)
    np.testing.assert_allclose(f32, f64, rtol=2e-5, atol=1e-5)


def test_max_sweeps_truncates_the_last_chunk(oracle):
    geom = GridGeometry.centered_cube(300.0, 20.0)
    field = np.zeros(geom.shape, dtype=np.float32)
    oracle.apply_dirichlet(field, DIRICHLET, 38.0)
    sweeps, res, conv = oracle.solve(field, np.zeros(geom.shape, np.float32), _args(oracle, geom, ob=38.0), 1.8, 25,
                                     1e-9, 10)
    assert sweeps == 25 and not conv and res > 1e-9


def test_colour_pass_touches_only_its_colour(oracle):
    geom = GridGeometry(origin=(0.0, 0.0, 0.0), dx=10.0, shape=(6, 5, 7))
    rng = np.random.default_rng(0)
    field = rng.uniform(5.0, 20.0, geom.shape).astype(np.float32)
    before = field.copy()
    a = _args(oracle, geom, boundary=(1, 1, 1))
    oracle.sor_sweep(field, rng.uniform(0, 1e-3, geom.shape).astype(np.float32), a, 1.5, 0)
    i, j, k = np.indices(geom.shape)
    black = (i + j + k) % 2 == 1
    np.testing.assert_array_equal(field[black], before[black])
    assert np.all(field[~black] != before[~black])


def test_uniform_production_source_equals_scaled_production(oracle):
    """A source species scales production by S/(K_S + S) (kernels/field_kernels.py:145-150): with a
    uniform source the solve equals the one whose production is pre-multiplied by that factor,
    evaluated as the kernel does (sf = S / (K_S + S), then production * sf, in float32)."""
    rng = np.random.default_rng(7)
    shape = (13, 11, 9)
    density = rng.uniform(0.0, 2e-3, shape).astype(np.float32)
    production = rng.uniform(0.0, 5.0, shape).astype(np.float32)
    S, K = np.float32(3.25), np.float32(0.75)
    sf = np.float32(S / np.float32(K + S))
    a = oracle.SolveArgs(dx=10.0, diffusion=7.2e5, uptake_max=2.0e6, michaelis_k=0.5, boundary_value=1.0)
    f1 = np.full(shape, 1.0, np.float32)
    f2 = f1.copy()
    r1 = oracle.solve(f1, density, a, 1.7, 400, 1e-6, 10, production=production,
                      source=np.full(shape, S, np.float32), source_k=float(K))
    r2 = oracle.solve(f2, density, a, 1.7, 400, 1e-6, 10, production=(production * sf).astype(np.float32))
    assert r1 == r2
    np.testing.assert_array_equal(f1, f2)
    # and the source matters: half the source concentration changes the field
    f3 = np.full(shape, 1.0, np.float32)
    oracle.solve(f3, density, a, 1.7, 400, 1e-6, 10, production=production,
                 source=np.full(shape, S / 2, np.float32), source_k=float(K))
    assert not np.array_equal(f1, f3)


def _residual_case(rng, shape, boundary, with_fixed, with_production, with_source):
    field = rng.uniform(0.0, 40.0, shape)
    field[rng.random(shape) < 0.05] = 0.0  # depleted nodes: the reference scale takes over
    density = np.where(rng.random(shape) < 0.5, rng.uniform(0.0, 2e-3, shape), 0.0)
    fixed = (rng.random(shape) < 0.15).astype(np.int32) if with_fixed else None
    production = rng.uniform(0.0, 5e3, shape) if with_production else None
    source = rng.uniform(0.0, 6.0, shape) if with_source else None
    return field, density, fixed, production, source


CASES = [((0, 0, 0), False, False, False), ((0, 1, 0), True, False, False), ((1, 1, 1), False, True, False),
         ((1, 0, 1), True, True, True), ((0, 0, 1), False, True, True)]


@pytest.mark.parametrize("boundary,with_fixed,with_production,with_source", CASES)
def test_residual_matches_float64_restatement(oracle, boundary, with_fixed, with_production, with_source):
    """The residual maximum against residual_terms_f64 (an independent vectorised restatement of
    kernels/field_kernels.py:215-243): 1e-12 relative in float64, within the float32 rounding of
    the terms' magnitudes in float32."""
    from tests.helpers import residual_terms_f64

    rng = np.random.default_rng(hash((boundary, with_fixed, with_production, with_source)) % 2 ** 32)
    shape = (9, 8, 10)
    field, density, fixed, production, source = _residual_case(rng, shape, boundary, with_fixed, with_production,
                                                               with_source)
    a = oracle.SolveArgs(dx=10.0, diffusion=7.2e4, uptake_max=5.5e5, michaelis_k=1.5, boundary_value=20.0,
                         boundary=boundary)
    terms, mask, scale = residual_terms_f64(field, density, a.dx2_over_D, a.uptake_max, a.michaelis_k, a.reference,
                                            boundary, fixed, production, source, 0.7)
    expected = terms[mask].max()
    got64 = oracle.residual(field, density, a, fixed=fixed, production=production, source=source, source_k=0.7)
    assert abs(got64 - expected) <= 1e-12 * expected
    f32 = lambda v: None if v is None else np.asarray(v, dtype=np.float32)  # noqa: E731
    got32 = oracle.residual(f32(field), f32(density), a, fixed=fixed, production=f32(production),
                            source=f32(source), source_k=0.7)
    terms32, _, _ = residual_terms_f64(f32(field), f32(density), a.dx2_over_D, a.uptake_max, a.michaelis_k,
                                       a.reference, boundary, fixed, f32(production), f32(source), 0.7)
    assert abs(got32 - terms32[mask].max()) <= 4e-7 * scale[mask].max()


def test_residual_term_of_every_node_matches_float64_restatement(oracle):
    """One free node at a time (every other node pinned): the oracle's residual is then that node's
    own term, so each node's term, including the mirrored faces, the reference scale at depleted
    nodes, production and its source factor, is compared and not only the maximum."""
    from tests.helpers import residual_terms_f64

    rng = np.random.default_rng(11)
    shape, boundary = (5, 4, 6), (1, 0, 1)
    field, density, _, production, source = _residual_case(rng, shape, boundary, False, True, True)
    a = oracle.SolveArgs(dx=10.0, diffusion=7.2e4, uptake_max=5.5e5, michaelis_k=1.5, boundary_value=20.0,
                         boundary=boundary)
    terms, mask, _ = residual_terms_f64(field, density, a.dx2_over_D, a.uptake_max, a.michaelis_k, a.reference,
                                        boundary, None, production, source, 0.7)
    checked = 0
    for t in np.ndindex(shape):
        fixed = np.ones(shape, dtype=np.int32)
        fixed[t] = 0
        got = oracle.residual(field, density, a, fixed=fixed, production=production, source=source, source_k=0.7)
        if mask[t]:
            assert abs(got - terms[t]) <= 1e-12 * max(terms[t], 1e-300), t
            checked += 1
        else:
            assert got == 0.0, t  # a node on a Dirichlet face is never part of the residual
    assert checked == int(mask.sum()) and checked > 0
