"""The metrics restatement (oracle.metrics_*, after metrics/population.py) pinned by closed forms,
and the host side of the device metrics (numpy's percentile interpolation) against numpy."""

import numpy as np
import pytest


def symmetric_cells(radii_by_shell):
    """Cells at +-r on each axis for every (radius, copies) pair: the float32 centroid is exactly 0
    and every centre distance is exactly its radius."""
    pts = []
    for r, copies in radii_by_shell:
        for _ in range(copies):
            for axis in range(3):
                for sign in (1.0, -1.0):
                    p = [0.0, 0.0, 0.0]
                    p[axis] = sign * r
                    pts.append(p)
    return np.array(pts, dtype=np.float32)


def test_state_counts_equal_bincount(oracle):
    s = np.random.default_rng(0).integers(0, 4, 1000).astype(np.int32)
    assert oracle.metrics_state_counts(s) == list(np.bincount(s, minlength=4))


def test_spheroid_radius_is_the_interpolated_99th_order_statistic(oracle):
    # 198 cells at radius 10 and 2 at radius 50: sorted distances [10 x 198, 50, 50]; the 99th
    # percentile sits at index 0.99 * 199 = 197.01, between the last 10 and the first 50.
    x = symmetric_cells([(10.0, 33), (50.0, 0)])
    x = np.vstack([x, np.array([[50, 0, 0], [-50, 0, 0]], dtype=np.float32)])
    n = x.shape[0]
    radii = np.full(n, 5.0, dtype=np.float32)
    states = np.zeros(n, dtype=np.int32)
    states[:7] = 3
    s = oracle.metrics_population_summary(x, radii, states)
    assert s["cells"] == 200 and s["dead_cells"] == 7 and s["living_cells"] == 193
    # numpy interpolates a float32 array in float32: 10 + 40 * gamma with gamma = 197.01 - 197
    # evaluated in float32
    gamma = np.float32(np.float32(199) * np.float32(0.99) - np.float32(197))
    assert s["spheroid_radius"] == float(np.float32(10) + np.float32(40) * gamma)
    assert s["spheroid_radius"] == pytest.approx(10.4, abs=1e-3)
    assert s["mean_radius"] == 5.0


def test_radial_profile_of_known_shells(oracle):
    x = symmetric_cells([(5.0, 2), (15.0, 1), (25.0, 3)])  # 12, 6 and 18 cells
    n = x.shape[0]
    r = np.linalg.norm(x, axis=1)
    oxygen = np.where(r < 10, 40.0, np.where(r < 20, 20.0, 5.0)).astype(np.float32)
    glucose = (oxygen / 10).astype(np.float32)
    states = np.where(r < 10, 3, np.where(r < 20, 2, 0)).astype(np.int32)
    p = oracle.metrics_radial_profile(x, oxygen, glucose, states, shell_width=10.0)
    assert list(p["r_outer"]) == [10.0, 20.0, 30.0]
    assert list(p["cells"]) == [12, 6, 18]
    assert list(p["oxygen_mean"]) == [40.0, 20.0, 5.0]
    assert list(p["glucose_mean"]) == pytest.approx([4.0, 2.0, 0.5])
    assert list(p["fraction_dead"]) == [1.0, 0.0, 0.0]
    assert list(p["fraction_hypoxic"]) == [0.0, 1.0, 0.0]
    assert list(p["fraction_proliferative"]) == [0.0, 0.0, 1.0]
    radii = oracle.metrics_shell_radii(p)
    assert radii == {"necrotic_radius": 10.0, "hypoxic_radius": 20.0, "non_proliferative_radius": 20.0}
    # equal bins up to the outermost cell: edges linspace(0, 25 + 1e-6, 6)
    q = oracle.metrics_radial_profile(x, oxygen, glucose, states, n_bins=5)
    assert q["cells"].sum() == n and q["cells"][-1] == 18
    assert np.isnan(q["oxygen_mean"][1]) and q["fraction_dead"][0] == 1.0


def test_oxygen_summary_over_interior_nodes_and_living_cells(oracle):
    f = np.arange(27, dtype=np.float32).reshape(3, 3, 3)
    interior = np.zeros((3, 3, 3), dtype=bool)
    interior[1, 1, 1] = True
    states = np.array([0, 3, 2], dtype=np.int32)
    o = np.array([4.0, -1.0, 2.0], dtype=np.float32)
    s = oracle.metrics_oxygen_summary(states, o, f, interior, 30)
    assert s == {"oxygen_cells_mean": 3.0, "oxygen_cells_min": 2.0, "oxygen_grid_mean": 13.0, "oxygen_grid_min": 13.0,
                 "field_sweeps": 30}


@pytest.mark.parametrize("n", [1, 2, 3, 100, 1001, 4096])
@pytest.mark.parametrize("percent", [99, 50, 0, 100, 37.5])
def test_host_percentile_interpolation_equals_numpy(n, percent):
    from warpbiocell_amd.metrics.population import _percentile

    a = np.random.default_rng(n).gamma(2.0, 50.0, n).astype(np.float32)
    srt = np.sort(a)
    assert _percentile(lambda k: srt[k], n, percent) == float(np.percentile(a, percent))
