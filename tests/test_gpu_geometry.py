"""GPU parity of the tissue-region path: the device node gradient against numpy.gradient, SDF
sampling, the confinement push (separate launch and fused into the contact gather) and confined
relaxation, bit for bit against the oracle; the reference's behavioural checks
(tests/test_geometry.py) restated on the GPU path; the tissue-surface oxygen boundary."""

import numpy as np
import pytest
import torch

from tests.helpers import steady_state_dense
from warpbiocell_amd.cells.initialization import spherical_cluster
from warpbiocell_amd.cells.mechanics import ContactParams, contact_substep, make_neighbor_grid, relax_contacts
from warpbiocell_amd.cells.state import CellPopulation
from warpbiocell_amd.fields.diffusion import SolverSettings
from warpbiocell_amd.fields.oxygen import OxygenField, OxygenParams
from warpbiocell_amd.fields.scalar_field import Boundary, GridGeometry
from warpbiocell_amd.geometry import Ellipsoid, Sphere, TissueRegion, node_gradient
from warpbiocell_amd.io.checkpoints import load_checkpoint, restore_region, save_checkpoint

pytestmark = pytest.mark.gpu

R = 8.0


def _oracle_region(oracle, region):
    return oracle.OracleRegion.from_sdf(region.sdf_numpy(), region.geometry.origin, region.geometry.dx)


@pytest.mark.parametrize("shape", [(41, 41, 41), (9, 8, 7), (2, 5, 3), (3, 2, 2)])
def test_node_gradient_is_bit_exact_against_numpy(device, shape):
    rng = np.random.default_rng(sum(shape))
    sdf = (rng.normal(size=shape) * 50.0).astype(np.float32)
    want = np.stack(np.gradient(sdf.astype(np.float64), 7.3, edge_order=1), axis=-1).astype(np.float32)
    got = node_gradient(sdf, 7.3, device).cpu().numpy()
    np.testing.assert_array_equal(got, want)


def test_region_gradient_and_sampling_are_bit_exact(oracle, device):
    geom = GridGeometry.centered_cube(400.0, 20.0)
    region = TissueRegion.from_shape(Ellipsoid(radii=(150.0, 120.0, 90.0)), geom, device=device)
    ref = _oracle_region(oracle, region)
    np.testing.assert_array_equal(region.gradient.cpu().numpy(), ref.gradient)
    rng = np.random.default_rng(3)
    pts = rng.uniform(-260.0, 260.0, size=(3000, 3)).astype(np.float32)  # some outside the grid
    pop = CellPopulation.from_numpy(pts, np.full(3000, R), capacity=3100, device=device)
    out = torch.zeros(3100, dtype=torch.float32, device=device)
    region.sample_at(pop.position, 3000, out)
    dev = out.cpu().numpy()[:3000]
    np.testing.assert_array_equal(dev, oracle.sample_sdf(ref, pts))
    np.testing.assert_allclose(dev, region.sample(pts), atol=1e-4)  # the float64 host query
    np.testing.assert_array_equal(region.contains(pts), region.sample(pts) < 0.0)


def test_region_sampling_agrees_with_the_analytical_distance(device):
    geom = GridGeometry.centered_cube(400.0, 20.0)
    region = TissueRegion.from_shape(Sphere(radius=150.0), geom, device=device)
    pts = np.random.default_rng(3).uniform(-180.0, 180.0, size=(300, 3)).astype(np.float32)
    host = region.sample(pts)
    assert np.abs(host - (np.linalg.norm(pts, axis=1) - 150.0)).max() < 0.15 * geom.dx
    assert abs(region.volume() / Sphere(radius=150.0).volume - 1.0) < 0.05


def _surface_cells(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pos = (d * rng.uniform(60.0, 140.0, (n, 1))).astype(np.float32)
    return pos, rng.uniform(6.0, 9.0, n).astype(np.float32)


def test_confinement_launch_is_bit_exact(oracle, device):
    """contact_substep with a region (separate gather, push and move launches)."""
    geom = GridGeometry.centered_cube(300.0, 20.0)
    region = TissueRegion.from_shape(Sphere(radius=100.0), geom, device=device)
    ref_region = _oracle_region(oracle, region)
    pos, radii = _surface_cells(4000, 7)
    params = ContactParams(stiffness=10.0, damping=1.0, max_radius=9.0, margin=0.5, wall_rate=12.0)
    pop = CellPopulation.from_numpy(pos, radii, device=device)
    contact_substep(pop, make_neighbor_grid(params, device), params, dt=0.01, region=region)
    q = params.query_radius
    vel, cnt = oracle.contact_velocities(pos, radii, np.float32(q), 128, q, params.rate)
    oracle.confinement_velocities(ref_region, params.wall_rate, pos, radii, vel)
    np.testing.assert_array_equal(pop.velocity.cpu().numpy()[:4000], vel)
    np.testing.assert_array_equal(pop.neighbor_counts_numpy(), cnt)
    np.testing.assert_array_equal(pop.positions_numpy(), pos + vel * np.float32(0.01))
    assert (np.abs(vel).sum(axis=1) > 0).sum() > 1000


def test_confined_relaxation_is_bit_exact_and_equals_separate_launches(oracle, device):
    geom = GridGeometry.centered_cube(300.0, 20.0)
    region = TissueRegion.from_shape(Sphere(radius=100.0), geom, device=device)
    pos, radii = _surface_cells(3000, 11)
    params = ContactParams(stiffness=10.0, damping=1.0, max_radius=9.0, margin=0.5, wall_rate=12.0)
    fused = CellPopulation.from_numpy(pos, radii, capacity=3100, device=device)
    relax_contacts(fused, make_neighbor_grid(params, device), params, dt=0.01, substeps=40, region=region)
    split = CellPopulation.from_numpy(pos, radii, capacity=3100, device=device)
    grid = make_neighbor_grid(params, device)
    for _ in range(40):
        contact_substep(split, grid, params, 0.01, region=region)
    ref = pos.copy()
    vel, cnt = np.zeros_like(ref), np.zeros(3000, np.int32)
    q = params.query_radius
    oracle.relax_contacts(ref, radii, np.float32(q), 128, q, params.rate, 0.01, 40, vel, cnt,
                          region=_oracle_region(oracle, region), wall_rate=params.wall_rate)
    np.testing.assert_array_equal(fused.positions_numpy(), ref)
    np.testing.assert_array_equal(fused.neighbor_counts_numpy(), cnt)
    np.testing.assert_array_equal(fused.velocity.cpu().numpy()[:3000], vel)
    np.testing.assert_array_equal(split.positions_numpy(), ref)
    assert np.abs(ref - pos).max() > 1.0


def test_confinement_pushes_cells_back_inside(device):
    geom = GridGeometry.centered_cube(400.0, 20.0)
    region = TissueRegion.from_shape(Sphere(radius=100.0), geom, device=device)
    pos = np.array([[130.0, 0.0, 0.0], [0.0, -120.0, 0.0], [0.0, 0.0, 0.0]], dtype=np.float32)
    pop = CellPopulation.from_numpy(pos, np.full(3, R), device=device)
    params = ContactParams(stiffness=10.0, damping=1.0, max_radius=R, wall_rate=10.0)
    relax_contacts(pop, make_neighbor_grid(params, device), params, dt=0.02, substeps=100, region=region)
    x = pop.positions_numpy()
    r = np.linalg.norm(x, axis=1)
    assert r[0] < 100.0 - R + 0.5 and r[1] < 100.0 - R + 0.5
    np.testing.assert_allclose(x[0] / r[0], [1.0, 0.0, 0.0], atol=1e-4)
    np.testing.assert_array_equal(x[2], [0.0, 0.0, 0.0])


def test_growing_population_stays_confined(oracle, device):
    geom = GridGeometry.centered_cube(400.0, 20.0)
    region = TissueRegion.from_shape(Sphere(radius=85.0), geom, device=device)
    pos = spherical_cluster(600, R, spacing_factor=0.85, jitter=0.2, seed=2)
    params = ContactParams(stiffness=10.0, damping=1.0, max_radius=R, wall_rate=10.0)
    pop = CellPopulation.from_numpy(pos, np.full(600, R), device=device)
    relax_contacts(pop, make_neighbor_grid(params, device), params, dt=0.02, substeps=300, region=region)
    r = np.linalg.norm(pop.positions_numpy(), axis=1)
    assert r.max() < 85.0
    free = CellPopulation.from_numpy(pos, np.full(600, R), device=device)
    relax_contacts(free, make_neighbor_grid(params, device), params, dt=0.02, substeps=300)
    assert np.linalg.norm(free.positions_numpy(), axis=1).max() > r.max() + 3.0
    ref = pos.copy()
    vel, cnt = np.zeros_like(ref), np.zeros(600, np.int32)
    oracle.relax_contacts(ref, np.full(600, R, np.float32), np.float32(2 * R), 128, 2 * R, 10.0, 0.02, 300, vel, cnt,
                          region=_oracle_region(oracle, region), wall_rate=10.0)
    np.testing.assert_array_equal(pop.positions_numpy(), ref)


def test_region_arguments_are_checked(device):
    geom = GridGeometry.centered_cube(100.0, 20.0)
    region = TissueRegion.from_shape(Sphere(radius=30.0), geom, device=device)
    with pytest.raises(ValueError, match="does not match"):
        TissueRegion(geom, np.zeros((3, 3, 3), np.float32), device=device)
    out = torch.zeros(4, dtype=torch.float32, device=device)
    with pytest.raises(ValueError):
        region.sample_at(torch.zeros((2, 3), dtype=torch.float32, device=device), 3, out)
    params = ContactParams(stiffness=1.0, damping=1.0, max_radius=R, wall_rate=40.0)
    pop = CellPopulation.from_numpy(np.zeros((1, 3), np.float32), np.full(1, R), device=device)
    with pytest.raises(ValueError, match="wall_rate"):
        relax_contacts(pop, make_neighbor_grid(params, device), params, dt=0.02, substeps=1, region=region)


def test_tissue_surface_oxygen_boundary(oracle, device):
    """fixed_outside pins every node outside the tissue to the boundary value: bit-exact against
    the oracle's solve with that mask, and within 2e-3 of a dense float64 solve."""
    geom = GridGeometry(origin=(0.0, 0.0, 0.0), dx=10.0, shape=(9, 8, 8))
    region = TissueRegion.from_shape(Sphere(center=(40.0, 35.0, 35.0), radius=28.0), geom, device=device)
    params = OxygenParams(diffusion_coefficient=2.0e3, uptake_max=5.0e4, michaelis_k=3.0, boundary_value=20.0)
    settings = SolverSettings(omega=1.5, tolerance=1e-5, max_sweeps=20000, check_every=10)
    ox = OxygenField(geom, params, settings, device=device, fixed_outside=region)
    density = np.where(region.inside_mask(), 1.5e-3, 0.0).astype(np.float32)
    ox.density.copy_(torch.from_numpy(density).to(device))
    report = ox.solve_steady_state()
    outside = ~region.inside_mask()
    values = ox.numpy()
    assert report.converged
    np.testing.assert_array_equal(values[outside], 20.0)
    assert values[region.inside_mask()].min() < 20.0
    a = oracle.SolveArgs(dx=geom.dx, diffusion=2.0e3, uptake_max=5.0e4, michaelis_k=3.0, boundary_value=20.0)
    ref = np.full(geom.shape, 20.0, np.float32)
    sweeps, res, conv = oracle.solve(ref, density, a, 1.5, 20000, 1e-5, 10, fixed=outside.astype(np.int32))
    np.testing.assert_array_equal(values, ref)
    assert (report.sweeps, np.float32(report.residual)) == (sweeps, np.float32(res))
    exact = steady_state_dense(np.full(geom.shape, 20.0), density, 2.0e3, 5.0e4, 3.0, geom.dx, (0, 0, 0),
                               fixed_extra=outside)
    np.testing.assert_allclose(values, exact, atol=2e-3)


def test_checkpoint_carries_the_region(device, tmp_path):
    geom = GridGeometry.centered_cube(200.0, 20.0)
    region = TissueRegion.from_shape(Sphere(radius=60.0), geom, device=device)
    pop = CellPopulation.from_numpy(spherical_cluster(50, R, seed=1), np.full(50, R), device=device)
    path = save_checkpoint(tmp_path / "c.npz", pop, None, 1.5, region=region)
    data = load_checkpoint(path)
    np.testing.assert_array_equal(data["tissue_sdf"], region.sdf_numpy())
    np.testing.assert_array_equal(data["grid_origin"], np.array(geom.origin))
    back = restore_region(data, device=device)
    np.testing.assert_array_equal(back.gradient.cpu().numpy(), region.gradient.cpu().numpy())
    assert back.geometry == geom
