"""The benchmark's pricing helpers, on the CPU: the HBM traffic it reads from the committed counter
table, and the count of quad lanes that load their density (the lanes inside the box of nonzero
density that the pack computes), against a lane-by-lane count."""

import numpy as np
import pytest
import torch

import bench


def test_the_committed_counter_table_prices_the_selected_shape():
    nodes = 161 ** 3
    sweep = bench.measured_traffic("packed/quad", "spheroid_100k_161", (161, 161, 161))
    assert sweep is not None
    # between the 12 B per node a quad sweep must move without any density and 20 B per node
    assert 12 * nodes <= sweep <= 20 * nodes
    assert bench.measured_traffic("packed/2", "spheroid_100k_161", (161, 161, 161)) is not None
    assert bench.measured_traffic("packed/quad", "spheroid_1m_257", (257, 257, 257)) is None
    assert bench.measured_traffic("packed/quad", "spheroid_100k_161", (385, 385, 385)) is None
    assert bench.measured_traffic("no/such/shape", "spheroid_100k_161", (161, 161, 161)) is None


def _box_lanes_by_lane(density: np.ndarray) -> int:
    """Lanes (plane i - 1, row j - 1, group q) of a quad pass over the interior planes and rows,
    q < the groups per row, whose coordinates lie inside the box of the nonzero densities'
    (i - 1, j - 1, k // 8), counted one by one."""
    nx, ny, nz = density.shape
    nonzero = np.argwhere(density != 0)  # NaN counts as nonzero
    if len(nonzero) == 0:
        return 0
    c = np.stack([nonzero[:, 0] - 1, nonzero[:, 1] - 1, nonzero[:, 2] // 8], axis=1)
    lo, hi = c.min(0), c.max(0)
    groups = (nz // 2 + 3) // 4
    count = 0
    for pl in range(nx - 2):
        for row in range(ny - 2):
            for q in range(groups):
                p = (pl, row, q)
                count += all(lo[a] <= p[a] <= hi[a] for a in range(3))
    return count


@pytest.mark.parametrize("shape", [(9, 12, 16), (17, 10, 66), (11, 13, 15)])
@pytest.mark.parametrize("pattern", ["empty", "corner", "faces", "cluster", "sparse", "nan"])
def test_quad_box_lanes_equal_a_lane_by_lane_count(shape, pattern):
    rng = np.random.default_rng(sum(shape) + len(pattern))
    d = np.zeros(shape, np.float32)
    nx, ny, nz = shape
    if pattern == "corner":
        d[1, 1, 1] = 1.0
    elif pattern == "faces":  # faces only: the box reaches past the interior lanes and is clipped
        d[0, :, :] = 1.0
        d[:, -1, 0] = 1.0
    elif pattern == "cluster":
        d[nx // 2 - 1:nx // 2 + 2, ny // 2 - 1:ny // 2 + 2, nz // 2 - 2:nz // 2 + 3] = 1.0
    elif pattern == "sparse":
        d[rng.random(shape) < 0.02] = 1.0
    elif pattern == "nan":
        d[nx - 2, 2, nz - 2] = np.nan
    got = bench.quad_box_lanes(torch.from_numpy(d), shape)
    assert got == _box_lanes_by_lane(d)
