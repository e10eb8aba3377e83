"""GPU parity of the hash-grid build and the contact mechanics: bucket order and neighbour
counts exact, velocities and positions bit-exact against the oracle; the reference's
brute-force and analytic checks restated on the GPU path."""

import numpy as np
import pytest

from tests.helpers import contact_brute_force
from warpbiocell_amd.cells.initialization import spherical_cluster
from warpbiocell_amd.cells.mechanics import ContactParams, contact_substep, make_neighbor_grid, relax_contacts
from warpbiocell_amd.cells.state import CellPopulation
from warpbiocell_amd.spatial.neighbors import NeighborGrid

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim", [5, 128])
def test_hash_grid_order_is_bit_exact(oracle, device, dim):
    pos = spherical_cluster(3000, 8.0, spacing_factor=0.85, jitter=0.3, seed=2)
    pop = CellPopulation.from_numpy(pos, np.full(3000, 8.0), capacity=3100, device=device)
    grid = NeighborGrid(16.5, dim=dim, device=device)
    for _ in range(2):  # the counts reset themselves: a second build must agree too
        grid.build(pop)
        start, ids = oracle.hashgrid_build(pos, 16.5, dim)
        np.testing.assert_array_equal(grid.point_ids().cpu().numpy(), ids)
        np.testing.assert_array_equal(grid.bucket_starts().cpu().numpy(), start)


def test_counts_match_brute_force_and_velocities_are_bit_exact(oracle, device):
    rng = np.random.default_rng(5)
    pos = spherical_cluster(2000, 8.0, spacing_factor=0.9, jitter=0.3, seed=5)
    radii = rng.uniform(5.0, 10.0, 2000).astype(np.float32)
    params = ContactParams(stiffness=1.0, damping=1.0, max_radius=10.0, margin=0.5)
    pop = CellPopulation.from_numpy(pos, radii, device=device)
    contact_substep(pop, make_neighbor_grid(params, device), params, dt=0.0)
    vel, cnt = oracle.contact_velocities(pos, radii, np.float32(params.query_radius), 128, params.query_radius,
                                         params.rate)
    v64, brute = contact_brute_force(pos, radii, params.rate, params.query_radius)
    np.testing.assert_array_equal(pop.neighbor_counts_numpy(), brute)
    np.testing.assert_array_equal(pop.neighbor_counts_numpy(), cnt)
    np.testing.assert_array_equal(pop.velocity.cpu().numpy(), vel)
    # against float64 (the accuracy bound for mechanics: 1e-5 relative)
    err = np.linalg.norm(pop.velocity.cpu().numpy() - v64) / np.linalg.norm(v64)
    print(f"contact velocities vs float64: relative L2 {err:.2e}")
    assert err < 1e-5
    np.testing.assert_array_equal(pop.positions_numpy(), pos)


def test_relaxation_substeps_are_bit_exact(oracle, device):
    r, n = 8.0, 3000
    pos = spherical_cluster(n, r, spacing_factor=0.85, jitter=0.2, seed=21)
    radii = np.random.default_rng(1).uniform(6.0, 9.0, n).astype(np.float32)
    params = ContactParams(stiffness=10.0, damping=1.0, max_radius=9.0, margin=0.5)
    pop = CellPopulation.from_numpy(pos, radii, capacity=n + 100, device=device)
    relax_contacts(pop, make_neighbor_grid(params, device), params, dt=0.01, substeps=40)
    ref = pos.copy()
    vel = np.zeros_like(ref)
    cnt = np.zeros(n, np.int32)
    oracle.relax_contacts(ref, radii, np.float32(params.query_radius), 128, params.query_radius, params.rate, 0.01, 40,
                          vel, cnt)
    np.testing.assert_array_equal(pop.positions_numpy(), ref)
    np.testing.assert_array_equal(pop.neighbor_counts_numpy(), cnt)
    assert np.abs(ref - pos).max() > 0.5


def test_substep_api_matches_the_fused_loop(device):
    pos = spherical_cluster(800, 8.0, spacing_factor=0.85, jitter=0.2, seed=7)
    params = ContactParams(stiffness=10.0, damping=1.0, max_radius=8.0, margin=0.5)
    a = CellPopulation.from_numpy(pos, np.full(800, 8.0), device=device)
    b = CellPopulation.from_numpy(pos, np.full(800, 8.0), device=device)
    ga, gb = make_neighbor_grid(params, device), make_neighbor_grid(params, device)
    for _ in range(10):
        contact_substep(a, ga, params, 0.01)
    relax_contacts(b, gb, params, 0.01, 10)
    np.testing.assert_array_equal(a.positions_numpy(), b.positions_numpy())
    np.testing.assert_array_equal(a.velocity.cpu().numpy(), b.velocity.cpu().numpy())


def test_grid_only_indexes_active_prefix(device):
    pos = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]], dtype=np.float32)
    params = ContactParams(stiffness=1.0, damping=1.0, max_radius=8.0)
    pop = CellPopulation.from_numpy(pos, np.full(2, 8.0), capacity=50, device=device)
    contact_substep(pop, make_neighbor_grid(params, device), params, dt=0.0)
    np.testing.assert_array_equal(pop.neighbor_counts_numpy(), [1, 1])


def test_isolated_cells_and_single_cell(device):
    pos = np.array([[0.0, 0.0, 0.0], [100.0, 0.0, 0.0], [0.0, 100.0, 0.0]], dtype=np.float32)
    params = ContactParams(stiffness=10.0, damping=1.0, max_radius=8.0)
    pop = CellPopulation.from_numpy(pos, np.full(3, 8.0), device=device)
    relax_contacts(pop, make_neighbor_grid(params, device), params, dt=0.01, substeps=50)
    np.testing.assert_array_equal(pop.positions_numpy(), pos)
    assert np.all(pop.neighbor_counts_numpy() == 0)


def test_pair_overlap_decays_at_the_analytic_rate(device):
    r, overlap0, dt, n = 8.0, 3.0, 0.01, 15
    pos = np.array([[0.0, 0.0, 0.0], [2 * r - overlap0, 0.0, 0.0]], dtype=np.float32)
    params = ContactParams(stiffness=10.0, damping=1.0, max_radius=r)
    pop = CellPopulation.from_numpy(pos, np.full(2, r), device=device)
    relax_contacts(pop, make_neighbor_grid(params, device), params, dt=dt, substeps=n)
    x = pop.positions_numpy()
    assert 2 * r - np.linalg.norm(x[1] - x[0]) == pytest.approx(overlap0 * (1 - 2 * params.rate * dt) ** n, rel=1e-3)
    assert np.all(x[:, 1:] == 0.0)


def test_runs_are_bitwise_reproducible(device):
    def run():
        pos = spherical_cluster(3000, 8.0, spacing_factor=0.85, jitter=0.2, seed=21)
        pop = CellPopulation.from_numpy(pos, np.full(3000, 8.0), device=device)
        params = ContactParams(stiffness=10.0, damping=1.0, max_radius=8.0, margin=0.5)
        relax_contacts(pop, make_neighbor_grid(params, device), params, dt=0.01, substeps=100)
        return pop.positions_numpy()

    np.testing.assert_array_equal(run(), run())


@pytest.mark.parametrize("factor,dim", [(0.45, 128), (0.7, 128), (1.3, 128), (2.5, 128), (0.7, 7), (1.0, 5),
                                        (0.3, 3), (0.2, 5)])
def test_other_bucket_sizes_take_the_general_walk(oracle, device, factor, dim):
    """Bucket edge != query radius: the walk's per-run path instead of the preloaded 3 x 3 rows
    (the (1.0, 5) case keeps the 3 x 3 rows with buckets wrapping in every axis; (0.3, 3) and
    (0.2, 5) have query boxes wider than the grid, cut to dim buckets per axis); counts and
    velocities equal the oracle's bits, and the counts equal brute force."""
    rng = np.random.default_rng(int(factor * 100) + dim)
    pos = spherical_cluster(1500, 8.0, spacing_factor=0.9, jitter=0.3, seed=int(factor * 10))
    radii = rng.uniform(5.0, 9.0, 1500).astype(np.float32)
    params = ContactParams(stiffness=1.0, damping=1.0, max_radius=9.0, margin=0.5)
    q = params.query_radius
    cell = np.float32(factor * q)
    pop = CellPopulation.from_numpy(pos, radii, device=device)
    contact_substep(pop, NeighborGrid(float(cell), dim=dim, device=device), params, dt=0.0)
    vel, cnt = oracle.contact_velocities(pos, radii, cell, dim, q, params.rate)
    np.testing.assert_array_equal(pop.neighbor_counts_numpy(), cnt)
    np.testing.assert_array_equal(pop.velocity.cpu().numpy()[:1500], vel)
    np.testing.assert_array_equal(cnt, contact_brute_force(pos, radii, params.rate, q)[1])


@pytest.mark.parametrize("dim", [3, 4, 6, 128])
def test_rows_whose_z_buckets_wrap_walk_as_two_runs(oracle, device, dim):
    """Bucket edge equal to the query radius, cluster centred on the origin: many cells' z search
    range crosses the dim boundary (for small dim, in every axis). The 3 x 3-row walk takes such a
    row as two runs. Counts, velocities and three moving substeps equal the oracle's bits."""
    pos = spherical_cluster(1500, 8.0, spacing_factor=0.9, jitter=0.3, seed=11)
    radii = np.random.default_rng(dim).uniform(5.0, 9.0, 1500).astype(np.float32)
    params = ContactParams(stiffness=1.0, damping=1.0, max_radius=9.0, margin=0.5)
    q = np.float32(params.query_radius)
    inv = np.float32(1.0) / q
    lo = np.floor((pos - q) * inv).astype(np.int64)
    hi = np.floor((pos + q) * inv).astype(np.int64)
    rows3x3 = (hi[:, 0] - lo[:, 0] == 2) & (hi[:, 1] - lo[:, 1] == 2) & (hi[:, 2] - lo[:, 2] < dim)
    z_wraps = np.mod(lo[:, 2], dim) + (hi[:, 2] - lo[:, 2]) >= dim
    assert (rows3x3 & z_wraps).sum() > 50
    pop = CellPopulation.from_numpy(pos, radii, device=device)
    contact_substep(pop, NeighborGrid(float(q), dim=dim, device=device), params, dt=0.0)
    vel, cnt = oracle.contact_velocities(pos, radii, q, dim, q, params.rate)
    np.testing.assert_array_equal(pop.neighbor_counts_numpy(), cnt)
    np.testing.assert_array_equal(pop.velocity.cpu().numpy()[:1500], vel)
    relax_contacts(pop, NeighborGrid(float(q), dim=dim, device=device), params, dt=0.05, substeps=3)
    ref, v3, c3 = pos.copy(), np.zeros_like(pos), np.zeros(1500, np.int32)
    oracle.relax_contacts(ref, radii, q, dim, q, params.rate, 0.05, 3, v3, c3)
    np.testing.assert_array_equal(pop.positions_numpy(), ref)
    np.testing.assert_array_equal(pop.neighbor_counts_numpy(), c3)


def test_dense_packing_drains_the_candidate_lists_in_order(oracle, device):
    """Over 100 neighbours inside the query radius per cell: every lane's candidate list fills and is
    drained several times per walk; the sums stay in walk order (bit-exact)."""
    pos = spherical_cluster(2500, 8.0, spacing_factor=0.3, jitter=0.3, seed=3)
    radii = np.full(2500, 8.0, np.float32)
    params = ContactParams(stiffness=1.0, damping=1.0, max_radius=8.0, margin=0.5)
    q = params.query_radius
    pop = CellPopulation.from_numpy(pos, radii, device=device)
    contact_substep(pop, make_neighbor_grid(params, device), params, dt=0.0)
    vel, cnt = oracle.contact_velocities(pos, radii, np.float32(q), 128, q, params.rate)
    assert np.median(cnt) > 100
    np.testing.assert_array_equal(pop.neighbor_counts_numpy(), cnt)
    np.testing.assert_array_equal(pop.velocity.cpu().numpy()[:2500], vel)
    relax_contacts(pop, make_neighbor_grid(params, device), params, dt=0.01, substeps=5)
    ref, v5, c5 = pos.copy(), np.zeros_like(pos), np.zeros(2500, np.int32)
    oracle.relax_contacts(ref, radii, np.float32(q), 128, q, params.rate, 0.01, 5, v5, c5)
    np.testing.assert_array_equal(pop.positions_numpy(), ref)


def test_relaxation_at_192_buckets_scans_its_tile_sums_in_a_launch(oracle, device):
    """dim = 192: 192^3 buckets are 1,728 scan tiles, beyond the 1,024 up to which each tile adds up
    the earlier tiles' sums itself, so every substep's bucket scan (the first from the tile-sum
    pass, later ones from the sums the gather added up) takes its tile-sum launch. Three substeps
    equal the oracle's bits."""
    r, n = 8.0, 3000
    pos = spherical_cluster(n, r, spacing_factor=0.85, jitter=0.2, seed=23)
    radii = np.random.default_rng(2).uniform(6.0, 9.0, n).astype(np.float32)
    params = ContactParams(stiffness=10.0, damping=1.0, max_radius=9.0, margin=0.5)
    q = np.float32(params.query_radius)
    pop = CellPopulation.from_numpy(pos, radii, device=device)
    relax_contacts(pop, NeighborGrid(float(q), dim=192, device=device), params, dt=0.01, substeps=3)
    ref, vel, cnt = pos.copy(), np.zeros_like(pos), np.zeros(n, np.int32)
    oracle.relax_contacts(ref, radii, q, 192, q, params.rate, 0.01, 3, vel, cnt)
    np.testing.assert_array_equal(pop.positions_numpy(), ref)
    np.testing.assert_array_equal(pop.neighbor_counts_numpy(), cnt)
    np.testing.assert_array_equal(pop.velocity.cpu().numpy()[:n], vel)
