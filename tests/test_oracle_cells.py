"""Pin the oracle's mechanics, hash grid, RNG, lifecycle and elementary functions against brute
force, closed forms and statistics (the checks of the reference's test_neighbors.py,
test_mechanics.py, test_lifecycle.py)."""

import math
from types import SimpleNamespace

import numpy as np
import pytest

from tests.helpers import contact_brute_force, pcg_py
from warpbiocell_amd.cells.initialization import spherical_cluster


def _lc(**kw):
    base = dict(death_rate=0.0, anoxic_death_rate=0.0, division_rate=0.0, inhibition_threshold=10 ** 6,
                hypoxia_threshold=0.0, death_threshold=0.0, glucose_threshold=0.0, glucose_death_threshold=0.0,
                necrosis_requires_glucose=False, phenotype_model="rules", apoptosis_rate=0.0, necrosis_rate=0.0,
                placement_factor=1.0)
    base.update(kw)
    return SimpleNamespace(**base)


# ---- elementary functions and RNG ---------------------------------------------------------------

def test_exp_within_two_ulp_of_the_exact_value(oracle):
    xs = np.concatenate([np.linspace(-30.0, 0.0, 4001), -np.logspace(-8, 1, 500)]).astype(np.float32)
    got = np.array([oracle.expf(float(x)) for x in xs], dtype=np.float64)
    ref = np.exp(xs.astype(np.float64))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got - ref) / ulp) <= 2.0
    assert oracle.expf(0.0) == 1.0 and oracle.expf(-200.0) == 0.0


def test_sin_cos_acos_accuracy(oracle):
    ts = np.linspace(0.0, 2 * math.pi, 4001).astype(np.float32)
    for f, ref in ((oracle.sinf, np.sin), (oracle.cosf, np.cos)):
        got = np.array([f(float(t)) for t in ts], dtype=np.float64)
        assert np.max(np.abs(got - ref(ts.astype(np.float64)))) < 1.5e-7
    us = np.linspace(-1.0, 1.0, 4001).astype(np.float32)
    got = np.array([oracle.acosf(float(u)) for u in us], dtype=np.float64)
    assert np.max(np.abs(got - np.arccos(us.astype(np.float64)))) < 4e-7
    assert oracle.acosf(1.0) == 0.0


def test_pcg_matches_an_independent_integer_restatement(oracle):
    rng = np.random.default_rng(0)
    for v in list(rng.integers(0, 2 ** 32, 2000)) + [0, 1, 2 ** 31, 2 ** 32 - 1]:
        assert oracle.pcg(int(v)) == pcg_py(int(v))


def test_rng_streams_are_keyed_on_slot_and_seed(oracle):
    a = oracle.init_rng(1000, 5)
    b = oracle.init_rng(3000, 5)
    np.testing.assert_array_equal(a, b[:1000])  # capacity does not change a slot's stream
    assert a[7] == pcg_py((5 + pcg_py(7)) % 2 ** 32)
    assert not np.array_equal(a, oracle.init_rng(1000, 6))
    assert len(np.unique(a)) > 990


def test_uniform_draws_have_the_right_moments(oracle):
    draws, _ = oracle.randf_sequence(12345, 20000)
    assert draws.min() >= 0.0 and draws.max() < 1.0
    assert abs(draws.mean() - 0.5) < 0.01 and abs(draws.var() - 1 / 12) < 0.005
    assert np.all(draws * 2 ** 24 == np.floor(draws * 2 ** 24))  # 24-bit grid


def test_sphere_directions_are_unit_and_isotropic(oracle):
    state = 99
    dirs = []
    for _ in range(4000):
        d, state = oracle.sample_sphere(state)
        dirs.append(d)
    dirs = np.array(dirs, dtype=np.float64)
    np.testing.assert_allclose(np.linalg.norm(dirs, axis=1), 1.0, atol=1e-6)
    assert np.all(np.abs(dirs.mean(axis=0)) < 0.05)
    np.testing.assert_allclose(np.cov(dirs.T), np.eye(3) / 3.0, atol=0.03)


# ---- hash grid and contact ----------------------------------------------------------------------

def test_hash_grid_is_a_stable_counting_sort(oracle):
    pos = spherical_cluster(500, 8.0, spacing_factor=0.85, jitter=0.3, seed=3)
    start, ids = oracle.hashgrid_build(pos, 16.5, 16)
    keys = np.array([oracle.lib().orc_bucket_key(*map(float, p), 1.0 / np.float32(16.5), 16) for p in pos])
    np.testing.assert_array_equal(ids, np.lexsort((np.arange(len(pos)), keys)))
    np.testing.assert_array_equal(np.diff(start), np.bincount(keys, minlength=16 ** 3))


def test_neighbor_counts_match_brute_force(oracle):
    r = 8.0
    pos = spherical_cluster(600, r, spacing_factor=0.85, jitter=0.3, seed=11)
    radii = np.full(600, r, dtype=np.float32)
    qr = 2 * r + 0.5
    vel, cnt = oracle.contact_velocities(pos, radii, qr, 128, qr, 1.0)
    v_ref, c_ref = contact_brute_force(pos, radii, 1.0, qr)
    np.testing.assert_array_equal(cnt, c_ref)
    np.testing.assert_allclose(vel, v_ref, atol=1e-4)
    assert c_ref.max() >= 6


def test_mixed_radii_and_aliasing_small_dim(oracle):
    """dim = 4 wraps many cells into each bucket; the neighbour sets must not change."""
    rng = np.random.default_rng(5)
    pos = spherical_cluster(400, 8.0, spacing_factor=1.0, jitter=0.2, seed=5)
    radii = rng.uniform(5.0, 10.0, size=400).astype(np.float32)
    for dim in (4, 128):
        vel, cnt = oracle.contact_velocities(pos, radii, 20.0, dim, 20.0, 10.0)
        v_ref, c_ref = contact_brute_force(pos, radii, 10.0, 20.0)
        np.testing.assert_array_equal(cnt, c_ref)
        np.testing.assert_allclose(vel, v_ref, atol=2e-3)


@pytest.mark.parametrize("factor,dim", [(0.3, 3), (0.45, 4), (0.2, 5)])
def test_query_box_wider_than_the_grid_counts_each_neighbour_once(oracle, factor, dim):
    """Buckets much smaller than the query radius on a small grid: the query box spans more than
    dim buckets per axis, so its coordinates wrap onto buckets already visited. Each neighbour
    is still counted once: counts equal brute force and velocities the float64 pair sums."""
    rng = np.random.default_rng(dim)
    pos = spherical_cluster(300, 8.0, spacing_factor=1.0, jitter=0.2, seed=dim)
    radii = rng.uniform(5.0, 10.0, size=300).astype(np.float32)
    q = 20.0
    assert 2 * q / (factor * q) + 1 > dim
    vel, cnt = oracle.contact_velocities(pos, radii, np.float32(factor * q), dim, q, 10.0)
    v_ref, c_ref = contact_brute_force(pos, radii, 10.0, q)
    np.testing.assert_array_equal(cnt, c_ref)
    np.testing.assert_allclose(vel, v_ref, atol=2e-3)


def test_pair_overlap_decays_at_the_analytic_rate(oracle):
    r, overlap0, dt, n = 8.0, 3.0, 0.01, 15
    pos = np.array([[0.0, 0.0, 0.0], [2 * r - overlap0, 0.0, 0.0]], dtype=np.float32)
    vel = np.zeros_like(pos)
    cnt = np.zeros(2, dtype=np.int32)
    oracle.relax_contacts(pos, np.full(2, r, np.float32), 2 * r, 128, 2 * r, 10.0, dt, n, vel, cnt)
    overlap = 2 * r - np.linalg.norm(pos[1] - pos[0])
    assert overlap == pytest.approx(overlap0 * (1 - 2 * 10.0 * dt) ** n, rel=1e-3)
    np.testing.assert_allclose(pos[0] + pos[1], [2 * r - overlap0, 0, 0], atol=1e-4)


def test_centre_of_mass_is_conserved(oracle):
    r = 8.0
    pos = spherical_cluster(1500, r, spacing_factor=0.85, jitter=0.2, seed=2)
    before = pos.astype(np.float64).mean(axis=0)
    vel = np.zeros_like(pos)
    oracle.relax_contacts(pos, np.full(1500, r, np.float32), 2 * r, 128, 2 * r, 10.0, 0.01, 100, vel,
                          np.zeros(1500, np.int32))
    np.testing.assert_allclose(pos.astype(np.float64).mean(axis=0), before, atol=1e-3)


def test_substeps_follow_the_pairwise_reference(oracle):
    r, n, dt = 8.0, 400, 0.01
    rng = np.random.default_rng(9)
    pos = spherical_cluster(n, r, spacing_factor=0.85, jitter=0.3, seed=9)
    radii = rng.uniform(6.0, 9.0, size=n).astype(np.float32)
    ref = pos.astype(np.float64)
    vel = np.zeros_like(pos)
    cnt = np.zeros(n, np.int32)
    for _ in range(20):
        oracle.relax_contacts(pos, radii, 18.5, 128, 18.5, 10.0, dt, 1, vel, cnt)
        v, _ = contact_brute_force(ref, radii, 10.0, 18.5)
        ref = ref + v * dt
        np.testing.assert_allclose(pos, ref, atol=2e-3)


# ---- scan and lifecycle -----------------------------------------------------------------------------

def test_scan_matches_cumsum(oracle):
    flags = (np.random.default_rng(1).random(10007) < 0.3).astype(np.int32)
    np.testing.assert_array_equal(oracle.scan_inclusive(flags), np.cumsum(flags))


def test_every_cell_divides_once_with_certain_division(oracle):
    n, R = 300, 8.0
    pos = spherical_cluster(n, R, spacing_factor=1.0, jitter=0.1, seed=0)
    pop = oracle.OraclePopulation.from_numpy(pos, np.full(n, R), capacity=3 * n)
    params = _lc(division_rate=1.0e3)
    oracle.lifecycle_decide(pop, params, 1.0)
    assert oracle.commit_divisions(pop, params) == n and pop.count == 2 * n
    x = pop.position[: 2 * n]
    np.testing.assert_allclose(np.linalg.norm(x[n:] - x[:n], axis=1), R, rtol=1e-5)
    np.testing.assert_allclose(0.5 * (x[n:] + x[:n]), pos, atol=1e-5)
    assert np.all(pop.age[:n] == 1.0) and np.all(pop.age[n: 2 * n] == 0.0)
    assert np.all(pop.cell_state[2 * n:] == 3) and np.all(pop.position[2 * n:] == 0.0)


def test_dead_cells_are_inert_and_death_is_certain(oracle):
    n = 200
    states = np.zeros(n, np.int32)
    states[::3] = 3
    pop = oracle.OraclePopulation.from_numpy(np.zeros((n, 3), np.float32) + np.arange(n)[:, None], np.full(n, 8.0),
                                             capacity=3 * n, states=states)
    rng_before = pop.rng_state.copy()
    oracle.lifecycle_decide(pop, _lc(division_rate=1e3), 1.0)
    dead = states == 3
    assert np.all(pop.age[:n][dead] == 0.0) and np.all(pop.divide_flag[:n][dead] == 0)
    np.testing.assert_array_equal(pop.rng_state[:n][dead], rng_before[:n][dead])
    pop2 = oracle.OraclePopulation.from_numpy(np.zeros((100, 3), np.float32), np.full(100, 8.0), capacity=300)
    oracle.lifecycle_decide(pop2, _lc(division_rate=1e3, death_rate=1e3), 1.0)
    assert np.all(pop2.cell_state[:100] == 3) and pop2.divide_flag[:100].sum() == 0


def test_death_rate_is_statistically_right(oracle):
    n, rate = 20000, 0.05
    pop = oracle.OraclePopulation.from_numpy(np.zeros((n, 3), np.float32), np.full(n, 8.0))
    oracle.lifecycle_decide(pop, _lc(death_rate=rate), 1.0)
    p = 1 - math.exp(-rate)
    assert abs((pop.cell_state == 3).mean() - p) < 4 * math.sqrt(p * (1 - p) / n)


def test_growth_follows_the_discrete_branching_law(oracle):
    n0, rate, steps = 2000, 0.1, 20
    pop = oracle.OraclePopulation.from_numpy(np.random.default_rng(0).uniform(-500, 500, (n0, 3)).astype(np.float32),
                                             np.full(n0, 8.0), capacity=25000)
    params = _lc(division_rate=rate)
    counts = [pop.count]
    for _ in range(steps):
        oracle.lifecycle_decide(pop, params, 1.0)
        oracle.commit_divisions(pop, params)
        counts.append(pop.count)
    p = 1 - math.exp(-rate)
    ratio = np.array(counts) / (n0 * (1 + p) ** np.arange(steps + 1))
    assert np.all(np.abs(ratio - 1) < 0.1) and abs(ratio[-1] - 1) < 0.07


def test_hypoxia_and_crowding_set_the_state(oracle):
    n = 6
    pop = oracle.OraclePopulation.from_numpy(np.zeros((n, 3), np.float32), np.full(n, 8.0), capacity=2 * n)
    pop.oxygen_local[:n] = [1.0, 5.0, 20.0, 20.0, 1.0, 20.0]
    pop.neighbor_count[:n] = [0, 0, 0, 9, 9, 9]
    oracle.lifecycle_decide(pop, _lc(hypoxia_threshold=8.0, death_threshold=2.0, inhibition_threshold=8), 0.1)
    np.testing.assert_array_equal(pop.cell_state[:n], [2, 2, 0, 1, 2, 1])


# ---- thread count ---------------------------------------------------------------------------------

def test_oracle_step_does_not_depend_on_the_thread_count(oracle):
    """The oracle's OpenMP loops (colour passes, residual maximum, contact gather, lifecycle) give
    the same bits on one thread and on eight: a full step with deaths, contacts and a solve."""
    from warpbiocell_amd import workloads as W

    w = W.spheroid_1k_41()
    a = oracle.SolveArgs(dx=w.dx_um, diffusion=W.D_UM2_PER_H, uptake_max=W.UPTAKE_MMHG_UM3_PER_H,
                         michaelis_k=W.MICHAELIS_K, boundary_value=w.boundary_mmHg)
    lifecycle = _lc(death_rate=2.0, hypoxia_threshold=20.0, death_threshold=5.0, anoxic_death_rate=1.0)
    contact = SimpleNamespace(query_radius=10.0, rate=10.0)
    stepping = SimpleNamespace(dt_cells=0.1, dt_mechanics=0.025, mechanics_substeps=4)
    results = []
    default_threads = oracle.get_threads()
    for threads in (1, 8):
        oracle.set_threads(threads)
        assert oracle.get_threads() == threads
        pop = oracle.OraclePopulation.from_numpy(w.positions, w.radii, seed=3)
        n = w.grid_n
        ox = oracle.OracleOxygen(shape=(n, n, n), origin=(-0.5 * (n - 1) * w.dx_um,) * 3, args=a, omega=1.8,
                                 max_sweeps=200, tolerance=1e-6, check_every=10)
        for _ in range(2):
            oracle.oracle_cell_step(pop, lifecycle, contact, stepping, oxygen=ox)
        results.append((pop, ox.values.copy()))
    oracle.set_threads(default_threads)
    (p1, f1), (p8, f8) = results
    assert np.array_equal(f1.view(np.uint32), f8.view(np.uint32))
    for name in ("position", "velocity", "neighbor_count", "cell_state", "age", "rng_state", "oxygen_local"):
        assert np.array_equal(getattr(p1, name).view(np.uint8), getattr(p8, name).view(np.uint8)), name
    assert (p1.cell_state[: p1.count] == 3).any()  # deaths happened
