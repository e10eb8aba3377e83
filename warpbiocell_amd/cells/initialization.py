"""Initial cell configurations (numpy only), as cells/initialization.py:8-38 of the reference.

``spherical_cluster`` keeps the ``n_cells`` points of a jittered cubic lattice closest to the
origin; the lattice spacing is ``spacing_factor * 2 * radius`` and the jitter is uniform in
``[-jitter*radius, jitter*radius]`` per coordinate, drawn from ``numpy.random.default_rng(seed)``.
"""

from __future__ import annotations

import numpy as np


def spherical_cluster(n_cells: int, radius: float, spacing_factor: float = 0.9, jitter: float = 0.1,
                      seed: int = 0) -> np.ndarray:
    if n_cells <= 0:
        raise ValueError("n_cells must be positive")
    rng = np.random.default_rng(seed)
    spacing = spacing_factor * 2.0 * radius
    # Half-width (in lattice steps) of a cube whose inscribed ball holds at least n_cells points.
    half = int(np.ceil(((n_cells / (np.pi / 6.0)) ** (1.0 / 3.0)) / 2.0)) + 1
    ticks = np.arange(-half, half + 1, dtype=np.float64) * spacing
    lattice = np.stack(np.meshgrid(ticks, ticks, ticks, indexing="ij"), axis=-1).reshape(-1, 3)
    keep = np.argsort(np.einsum("ij,ij->i", lattice, lattice), kind="stable")[:n_cells]
    points = lattice[keep]
    points += rng.uniform(-jitter * radius, jitter * radius, size=points.shape)
    return points.astype(np.float32)
