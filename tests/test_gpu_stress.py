"""The stress configuration at full size: 1,000,000 cells on a 257^3 grid with 5 % of the cells
dividing per step (prefix-sum slot allocation at scale), fixed 20 red-black sweeps per step.

The first step is compared bit for bit with the oracle (OpenMP on the host cores). All 20 steps
then run on the GPU, and the population must grow as the branching law says: every living cell
divides with probability 0.05 per step and none dies, so the count after 20 steps has mean
n0 * 1.05^20 and a standard deviation of a few hundred cells.
"""

import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ARRAYS = ("position", "radius", "cell_state", "age", "oxygen_local", "velocity", "neighbor_count", "divide_flag",
          "division_offset")


def test_stress_step_is_bit_exact_and_growth_follows_the_branching_law(oracle, device):
    from warpbiocell_amd import workloads as W
    from warpbiocell_amd.simulation import scenario
    from warpbiocell_amd.simulation.simulator import cell_step

    w = W.spheroid_1m_257()
    sc = scenario.build(w, device)
    n0 = sc.population.count
    assert n0 == 1_000_000 and sc.geometry.shape == (257, 257, 257)

    ref = oracle.OraclePopulation.from_numpy(w.positions, w.radii, capacity=w.capacity, seed=w.seed)
    a = oracle.SolveArgs(dx=sc.geometry.dx, diffusion=sc.params.diffusion_coefficient,
                         uptake_max=sc.params.uptake_max, michaelis_k=sc.params.michaelis_k,
                         boundary_value=sc.params.boundary_value)
    rox = oracle.OracleOxygen(shape=sc.geometry.shape, origin=sc.geometry.origin, args=a, omega=w.omega,
                              max_sweeps=w.max_sweeps, tolerance=w.tolerance, check_every=w.check_every)

    print("stress: step 0 on the GPU", flush=True)
    report = cell_step(sc.population, sc.grid, sc.lifecycle, sc.contact, sc.stepping, oxygen=sc.oxygen)
    print("stress: step 0 on the oracle", flush=True)
    daughters, field = oracle.oracle_cell_step(ref, sc.lifecycle, sc.contact, sc.stepping, oxygen=rox)
    assert report.field_sweeps == field[0] == 20
    assert report.daughters == daughters > 0
    assert sc.population.count == ref.count == n0 + daughters
    n = ref.count
    for name in ARRAYS:
        got = getattr(sc.population, name)[:n].cpu().numpy()
        assert np.array_equal(got.view(np.uint8), getattr(ref, name)[:n].view(np.uint8)), name
    assert np.array_equal(sc.population.rng_numpy()[:n], ref.rng_state[:n])
    assert np.array_equal(sc.oxygen.numpy().view(np.uint32), rox.values.view(np.uint32))

    counts = [sc.population.count]
    for step in range(1, 20):
        report = cell_step(sc.population, sc.grid, sc.lifecycle, sc.contact, sc.stepping, oxygen=sc.oxygen)
        assert report.field_sweeps == 20
        counts.append(sc.population.count)
    print("stress: counts", counts[::5], counts[-1], flush=True)
    # Each step multiplies the expected count by 1 + p with p = 1 - exp(-rate dt) = 0.05 (the
    # lifecycle's own probability, evaluated in float); the variance of a Galton-Watson process
    # with Bernoulli(p) offspring is n0 p(1-p) m^(t-1) (m^t - 1)/(m - 1) with m = 1 + p.
    p = 1.0 - math.exp(-sc.lifecycle.division_rate * sc.stepping.dt_cells)
    m, t = 1.0 + p, 20
    mean = n0 * m ** t
    sd = math.sqrt(n0 * p * (1 - p) * m ** (t - 1) * (m ** t - 1) / (m - 1))
    assert abs(counts[-1] - mean) < 6 * sd, (counts[-1], mean, sd)
    assert counts[-1] <= w.capacity
    pos = sc.population.position[: counts[-1]].cpu().numpy()
    assert np.isfinite(pos).all()
    assert np.isfinite(sc.oxygen.numpy()).all() and sc.oxygen.numpy().min() >= 0.0
