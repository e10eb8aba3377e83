"""Pin the oracle's tissue-region functions (node gradient, SDF sampling, confinement push,
confined relaxation) against closed forms, a float64 restatement of the push written from the
confinement law, and the reference's behavioural checks (tests/test_geometry.py), and the host
shapes against their analytic distances."""

import numpy as np
import pytest

from warpbiocell_amd.cells.initialization import spherical_cluster
from warpbiocell_amd.fields.scalar_field import GridGeometry
from warpbiocell_amd.geometry.shapes import Ellipsoid, Sphere, Union, evaluate_on_grid

R = 8.0


def _region(O, shape, geom):
    return O.OracleRegion.from_sdf(evaluate_on_grid(shape, geom), geom.origin, geom.dx)


def test_sphere_sdf_is_exact_on_nodes():
    geom = GridGeometry.centered_cube(400.0, 20.0)
    sdf = evaluate_on_grid(Sphere(radius=150.0), geom)
    x, y, z = np.meshgrid(*geom.node_coordinates(), indexing="ij")
    np.testing.assert_allclose(sdf, np.sqrt(x ** 2 + y ** 2 + z ** 2) - 150.0, atol=1e-4)
    volume = float((sdf < 0).sum()) * geom.voxel_volume
    assert abs(volume / Sphere(radius=150.0).volume - 1.0) < 0.05


def test_ellipsoid_and_union_have_the_right_sign_and_bounds():
    ell = Ellipsoid(center=(10.0, 0.0, 0.0), radii=(200.0, 100.0, 50.0))
    d = ell.sdf(np.array([10.0, 209.0, 211.0, 10.0, 10.0]), np.array([0.0, 0.0, 0.0, 99.0, 101.0]), np.zeros(5))
    assert d[0] < 0 and d[1] < 0 and d[2] > 0 and d[3] < 0 and d[4] > 0
    assert abs(d[1] + 1.0) < 0.2 and abs(d[2] - 1.0) < 0.2
    lo, hi = ell.bounds
    np.testing.assert_allclose(lo, [-190.0, -100.0, -50.0])
    np.testing.assert_allclose(hi, [210.0, 100.0, 50.0])
    union = Union((Sphere(center=(-100.0, 0.0, 0.0), radius=80.0), Sphere(center=(100.0, 0.0, 0.0), radius=80.0)))
    d = union.sdf(np.array([-100.0, 100.0, 0.0]), np.zeros(3), np.zeros(3))
    assert d[0] == -80.0 and d[1] == -80.0 and d[2] == 20.0
    with pytest.raises(ValueError):
        Union(())
    with pytest.raises(ValueError):
        Sphere(radius=0.0)


def test_node_gradient_is_exact_on_linear_fields_and_matches_the_difference_stencil(oracle):
    shape, dx = (7, 5, 4), 2.5
    i, j, k = np.meshgrid(*(np.arange(n) for n in shape), indexing="ij")
    linear = (3.0 * i - 2.0 * j + 0.5 * k).astype(np.float32) * np.float32(dx)
    g = oracle.node_gradient(linear, dx)
    np.testing.assert_array_equal(g[..., 0], 3.0)
    np.testing.assert_array_equal(g[..., 1], -2.0)
    np.testing.assert_array_equal(g[..., 2], 0.5)
    # A quadratic: second-order central differences are exact inside, first order on the faces.
    rng = np.random.default_rng(0)
    f = rng.normal(size=shape).astype(np.float32)
    g = oracle.node_gradient(f, dx)
    f64 = f.astype(np.float64)
    for axis in range(3):
        n = shape[axis]
        for idx in range(n):
            lo, hi = max(idx - 1, 0), min(idx + 1, n - 1)
            want = (np.take(f64, hi, axis) - np.take(f64, lo, axis)) / ((hi - lo) * dx)
            np.testing.assert_array_equal(np.take(g[..., axis], idx, axis), want.astype(np.float32))


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_planar_wall_push_has_the_closed_form(oracle, axis):
    """SDF = coordinate - 60 along one axis: |grad| = 1, so the push is -wall_rate * (sdf + r)
    along that axis only, zero for cells whose surface stays inside, and the other components
    are untouched bit for bit."""
    geom = GridGeometry(origin=(0.0, 0.0, 0.0), dx=10.0, shape=(13, 11, 9))
    coords = np.meshgrid(*geom.node_coordinates(), indexing="ij")
    sdf = (coords[axis] - 60.0).astype(np.float32)
    region = oracle.OracleRegion.from_sdf(sdf, geom.origin, geom.dx)
    rng = np.random.default_rng(axis)
    n = 400
    pos = rng.uniform(5.0, 75.0, size=(n, 3)).astype(np.float32)
    radii = rng.uniform(4.0, 9.0, n).astype(np.float32)
    v0 = rng.normal(size=(n, 3)).astype(np.float32)
    vel = v0.copy()
    oracle.confinement_velocities(region, 10.0, pos, radii, vel)
    p = pos[:, axis].astype(np.float64) - 60.0 + radii
    out = p > 0
    want = v0[:, axis].astype(np.float64) - np.where(out, 10.0 * p, 0.0)
    np.testing.assert_allclose(vel[:, axis], want, rtol=0, atol=2e-5 * (1 + np.abs(want)).max())
    np.testing.assert_array_equal(vel[~out], v0[~out])
    others = [a for a in range(3) if a != axis]
    np.testing.assert_array_equal(vel[:, others], v0[:, others])
    assert 50 < out.sum() < n - 50


def test_sphere_push_matches_a_float64_restatement(oracle):
    """Trilinear SDF and gradient at each cell, then -wall_rate * p / |g| * g where p > 0 and
    |g| >= 1e-8, in float64 from the law; the float32 oracle agrees to float32 rounding. Cells
    outside the grid are clamped onto it."""
    geom = GridGeometry.centered_cube(300.0, 20.0)
    region = _region(oracle, Sphere(radius=100.0), geom)
    rng = np.random.default_rng(3)
    pos = rng.uniform(-165.0, 165.0, size=(1500, 3)).astype(np.float32)
    radii = rng.uniform(5.0, 10.0, 1500).astype(np.float32)
    vel = np.zeros_like(pos)
    oracle.confinement_velocities(region, 7.5, pos, radii, vel)

    shape = np.array(geom.shape)
    u = np.clip((pos.astype(np.float64) - np.asarray(geom.origin)) / geom.dx, 0.0, shape - 1.0)
    base = np.minimum(np.floor(u).astype(int), shape - 2)
    f = u - base
    weights, corners = [], []
    for di in (0, 1):
        for dj in (0, 1):
            for dk in (0, 1):
                w = (f[:, 0] if di else 1 - f[:, 0]) * (f[:, 1] if dj else 1 - f[:, 1]) * (f[:, 2] if dk else 1 - f[:, 2])
                weights.append(w)
                corners.append((base[:, 0] + di, base[:, 1] + dj, base[:, 2] + dk))
    sdf = sum(w * region.sdf.astype(np.float64)[c] for w, c in zip(weights, corners))
    grad = sum(w[:, None] * region.gradient.astype(np.float64)[c] for w, c in zip(weights, corners))
    p = sdf + radii
    length = np.linalg.norm(grad, axis=1)
    push = np.where((p > 0) & (length >= 1e-8), 7.5 * p / np.where(length > 0, length, 1.0), 0.0)
    want = -push[:, None] * grad
    np.testing.assert_allclose(vel, want, rtol=0, atol=1e-4 * max(1.0, np.abs(want).max()))
    assert (push > 0).sum() > 300 and (push == 0).sum() > 80


def test_degenerate_gradient_and_empty_input_leave_velocities_alone(oracle):
    geom = GridGeometry(origin=(0.0, 0.0, 0.0), dx=10.0, shape=(5, 5, 5))
    region = oracle.OracleRegion.from_sdf(np.full(geom.shape, 5.0, np.float32), geom.origin, geom.dx)
    pos = np.array([[20.0, 20.0, 20.0], [-50.0, 90.0, 10.0]], np.float32)
    vel = np.array([[1.0, -2.0, 3.0], [0.5, 0.25, -1.0]], np.float32)
    v0 = vel.copy()
    oracle.confinement_velocities(region, 10.0, pos, np.full(2, 8.0, np.float32), vel)
    np.testing.assert_array_equal(vel, v0)  # p = 13 > 0 but grad = 0
    empty = np.zeros((0, 3), np.float32)
    oracle.confinement_velocities(region, 10.0, empty, np.zeros(0, np.float32), empty)


def test_sample_sdf_matches_the_analytic_distance(oracle):
    geom = GridGeometry.centered_cube(400.0, 20.0)
    region = _region(oracle, Sphere(radius=150.0), geom)
    pts = np.random.default_rng(3).uniform(-180.0, 180.0, size=(300, 3)).astype(np.float32)
    s = oracle.sample_sdf(region, pts)
    assert np.abs(s - (np.linalg.norm(pts, axis=1) - 150.0)).max() < 0.15 * geom.dx


def test_confinement_pushes_cells_back_inside(oracle):
    geom = GridGeometry.centered_cube(400.0, 20.0)
    region = _region(oracle, Sphere(radius=100.0), geom)
    pos = np.array([[130.0, 0.0, 0.0], [0.0, -120.0, 0.0], [0.0, 0.0, 0.0]], dtype=np.float32)
    radii = np.full(3, R, np.float32)
    vel, cnt = np.zeros_like(pos), np.zeros(3, np.int32)
    q = 2 * R
    oracle.relax_contacts(pos, radii, np.float32(q), 128, q, 10.0, 0.02, 100, vel, cnt, region=region, wall_rate=10.0)
    r = np.linalg.norm(pos, axis=1)
    assert r[0] < 100.0 - R + 0.5 and r[1] < 100.0 - R + 0.5
    np.testing.assert_allclose(pos[0] / r[0], [1.0, 0.0, 0.0], atol=1e-4)
    np.testing.assert_array_equal(pos[2], [0.0, 0.0, 0.0])


def test_growing_population_stays_confined(oracle):
    geom = GridGeometry.centered_cube(400.0, 20.0)
    region = _region(oracle, Sphere(radius=85.0), geom)
    start = spherical_cluster(600, R, spacing_factor=0.85, jitter=0.2, seed=2)
    radii = np.full(600, R, np.float32)
    q = 2 * R
    confined, free = start.copy(), start.copy()
    for pos, reg in ((confined, region), (free, None)):
        vel, cnt = np.zeros_like(pos), np.zeros(600, np.int32)
        oracle.relax_contacts(pos, radii, np.float32(q), 128, q, 10.0, 0.02, 300, vel, cnt, region=reg, wall_rate=10.0)
    r = np.linalg.norm(confined, axis=1)
    assert r.max() < 85.0
    assert np.linalg.norm(free, axis=1).max() > r.max() + 3.0


def test_pinned_nodes_outside_the_tissue_match_a_dense_solve(oracle):
    """Pinned (``fixed``) nodes in the oracle's SOR: the reference's tissue-surface boundary case
    (tests/test_geometry.py:136-153) against a dense float64 solve with the same nodes held."""
    from tests.helpers import steady_state_dense

    geom = GridGeometry(origin=(0.0, 0.0, 0.0), dx=10.0, shape=(9, 8, 8))
    sdf = evaluate_on_grid(Sphere(center=(40.0, 35.0, 35.0), radius=28.0), geom)
    inside = sdf < 0.0
    density = np.where(inside, 1.5e-3, 0.0).astype(np.float32)
    a = oracle.SolveArgs(dx=geom.dx, diffusion=2.0e3, uptake_max=5.0e4, michaelis_k=3.0, boundary_value=20.0)
    values = np.full(geom.shape, 20.0, np.float32)
    sweeps, res, conv = oracle.solve(values, density, a, 1.5, 20000, 1e-5, 10, fixed=(~inside).astype(np.int32))
    assert conv and res < 1e-5
    np.testing.assert_array_equal(values[~inside], 20.0)
    assert values[inside].min() < 20.0
    exact = steady_state_dense(np.full(geom.shape, 20.0), density, 2.0e3, 5.0e4, 3.0, geom.dx, (0, 0, 0),
                               fixed_extra=~inside)
    np.testing.assert_allclose(values, exact, atol=2e-3)
    free = np.full(geom.shape, 20.0, np.float32)
    oracle.solve(free, density, a, 1.5, 20000, 1e-5, 10)
    assert np.abs(free - exact).max() > 0.05  # the pinning changes the answer
