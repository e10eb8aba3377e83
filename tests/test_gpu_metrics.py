"""Device-side population metrics (warpbiocell_amd.metrics.population) against the restatement of
the reference's host-side numpy metrics (oracle.metrics_*, after metrics/population.py).

Counts and minima are compared exactly; sums are double on the device where numpy accumulates in
float32, so means are compared at 1e-6 relative (FP32 tolerance). The device centre is the exact
mean rounded to float32 (checked within one ulp) and its centre distances are numpy's float32 norm
about that centre bit for bit; the radial profile's shells are then compared exactly with the
restatement binning those distances.
"""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def make_population(device, n, capacity, seed):
    from warpbiocell_amd.cells.state import CellPopulation

    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True) + 1e-12
    x = (d * (240.0 * rng.random(n) ** (1 / 3))[:, None] + np.array([3.0, -7.0, 11.0])).astype(np.float32)
    radii = (5.0 * np.exp(0.1 * rng.standard_normal(n))).astype(np.float32)
    pop = CellPopulation.from_numpy(x, radii, capacity=capacity, device=device, seed=seed)
    if n:
        states = rng.integers(0, 4, n).astype(np.int32)
        pop.cell_state[:n] = torch.from_numpy(states).to(device)
        pop.oxygen_local[:n] = torch.from_numpy(rng.uniform(0.0, 60.0, n).astype(np.float32)).to(device)
        pop.glucose_local[:n] = torch.from_numpy(rng.uniform(0.0, 5.0, n).astype(np.float32)).to(device)
    torch.cuda.synchronize()
    return pop


def make_oxygen(device, seed, boundary, pinned):
    from warpbiocell_amd import workloads as W
    from warpbiocell_amd.fields.diffusion import SolverSettings
    from warpbiocell_amd.fields.oxygen import OxygenField, OxygenParams
    from warpbiocell_amd.fields.scalar_field import GridGeometry

    geom = GridGeometry.centered_cube(600.0, 20.0)
    ox = OxygenField(geom, OxygenParams(W.D_UM2_PER_H, W.UPTAKE_MMHG_UM3_PER_H, W.MICHAELIS_K, 40.0),
                     SolverSettings(omega=1.8, tolerance=1e-6, max_sweeps=2000, check_every=10), boundary=boundary,
                     device=device)
    rng = np.random.default_rng(seed)
    ox.field.assign(rng.uniform(1.0, 40.0, geom.shape).astype(np.float32))
    if pinned:
        ox.field.set_fixed(rng.random(geom.shape) < 0.05)
    torch.cuda.synchronize()
    return ox


def host(pop):
    n = pop.count
    return (pop.position[:n].cpu().numpy(), pop.radius[:n].cpu().numpy(), pop.cell_state[:n].cpu().numpy(),
            pop.oxygen_local[:n].cpu().numpy(), pop.glucose_local[:n].cpu().numpy())


def close(a, b, rel=1e-6):
    if np.isnan(a) or np.isnan(b):
        return bool(np.isnan(a) and np.isnan(b))
    return abs(a - b) <= rel * max(abs(a), abs(b), 1e-30)


def same(a, b):
    """Equal, NaN included (no living cell gives NaN on both sides)."""
    if isinstance(a, float) and isinstance(b, float) and np.isnan(a) and np.isnan(b):
        return True
    return a == b


@pytest.mark.parametrize("n, capacity, pinned, neumann", [(20000, 24000, False, False), (777, 777, True, True),
                                                           (1, 4, False, False)])
def test_population_summary_matches_the_numpy_metrics(device, oracle, n, capacity, pinned, neumann):
    from warpbiocell_amd.fields.scalar_field import Boundary
    from warpbiocell_amd.metrics.population import population_summary, state_counts

    pop = make_population(device, n, capacity, seed=n)
    boundary = (Boundary.DIRICHLET, Boundary.NEUMANN if neumann else Boundary.DIRICHLET, Boundary.DIRICHLET)
    ox = make_oxygen(device, n, boundary, pinned)
    got = population_summary(pop, ox)
    x, radii, states, o, _ = host(pop)
    ref = oracle.metrics_population_summary(x, radii, states)
    ref.update(oracle.metrics_oxygen_summary(states, o, ox.field.values.cpu().numpy(), ox.field.interior_mask(), 0))
    for key in ("cells", "living_cells", "dead_cells", "proliferative_cells", "quiescent_cells", "hypoxic_cells",
                "oxygen_cells_min", "oxygen_grid_min", "field_sweeps"):
        assert same(got[key], ref[key]), (key, got[key], ref[key])
    for key in ("spheroid_radius", "mean_radius", "oxygen_cells_mean", "oxygen_grid_mean"):
        assert close(got[key], ref[key]), (key, got[key], ref[key])
    assert {int(k): v for k, v in state_counts(pop).items()} == dict(enumerate(oracle.metrics_state_counts(states)))


def test_empty_population(device, oracle):
    from warpbiocell_amd.metrics.population import population_summary, radial_profile

    pop = make_population(device, 0, 8, seed=1)
    ox = make_oxygen(device, 1, (0, 0, 0), False)
    got = population_summary(pop, ox)
    assert got["cells"] == 0 and got["spheroid_radius"] == 0.0 and got["mean_radius"] == 0.0
    assert np.isnan(got["oxygen_cells_mean"]) and np.isnan(got["oxygen_cells_min"])
    p = radial_profile(pop, n_bins=4)
    assert list(p["cells"]) == [0, 0, 0, 0] and np.isnan(p["oxygen_mean"]).all()


@pytest.mark.parametrize("kw", [dict(n_bins=10), dict(n_bins=None), dict(shell_width=25.0), dict(n_bins=1)])
def test_radial_profile_matches_the_numpy_metrics(device, oracle, kw):
    from warpbiocell_amd.metrics import population as M

    pop = make_population(device, 20000, 20000, seed=3)
    pop.local_field("lactate")[:] = torch.linspace(0.0, 9.0, pop.capacity, device=device)
    got = M.radial_profile(pop, **kw)
    x, _, states, o, g = host(pop)
    lac = pop.extra_local["lactate"][: pop.count].cpu().numpy()
    ws = M._workspace(pop)
    m = M._cell_moments(pop, ws)
    r_dev, _ = M._centre_distances(pop, m, ws)
    r_dev = r_dev.cpu().numpy()
    # the device centre is the exact mean rounded to float32 (numpy's float32 running sum drifts);
    # about that centre the distances are numpy's float32 norm bit for bit
    c_dev = M._centroid(m)
    exact = x.astype(np.float64).mean(axis=0)
    assert np.all(np.abs(c_dev - exact) <= np.spacing(np.abs(exact).astype(np.float32)))
    np.testing.assert_array_equal(r_dev, np.linalg.norm(x - c_dev, axis=1))
    ref = oracle.metrics_radial_profile(x, o, g, states, lactate_local=lac, r=r_dev, **kw)
    assert set(got) == set(ref)
    np.testing.assert_array_equal(got["r_outer"], ref["r_outer"])
    np.testing.assert_array_equal(got["cells"], ref["cells"])
    for k in ("fraction_proliferative", "fraction_quiescent", "fraction_hypoxic", "fraction_dead"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    for k in ("oxygen_mean", "glucose_mean", "lactate_mean"):
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-6, err_msg=k)
    assert M.shell_radii(got) == oracle.metrics_shell_radii(ref)


def test_metrics_repeat_bit_for_bit(device):
    from warpbiocell_amd.metrics.population import population_summary, radial_profile

    pop = make_population(device, 50000, 50000, seed=5)
    ox = make_oxygen(device, 5, (0, 0, 0), True)
    a, b = population_summary(pop, ox), population_summary(pop, ox)
    assert a == b
    p, q = radial_profile(pop, n_bins=12), radial_profile(pop, n_bins=12)
    for k in p:
        np.testing.assert_array_equal(p[k], q[k])
