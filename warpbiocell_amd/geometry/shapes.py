"""Synthetic tissue shapes as signed-distance functions, negative inside, evaluated on grid nodes
(geometry/shapes.py:17-92 of the reference).

Lengths in um. The sphere and the union are exact distances; the ellipsoid uses the first-order
estimate |f| / |grad f| of its implicit function, exact on the axes.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from warpbiocell_amd.fields.scalar_field import GridGeometry


@dataclass(frozen=True)
class Sphere:
    center: tuple[float, float, float] = (0.0, 0.0, 0.0)
    radius: float = 200.0

    def __post_init__(self):
        if self.radius <= 0.0:
            raise ValueError("radius must be positive")

    def sdf(self, x, y, z) -> np.ndarray:
        cx, cy, cz = self.center
        return np.sqrt((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) - self.radius

    @property
    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        c = np.asarray(self.center, dtype=float)
        return c - self.radius, c + self.radius

    @property
    def volume(self) -> float:
        return 4.0 / 3.0 * np.pi * self.radius ** 3


@dataclass(frozen=True)
class Ellipsoid:
    center: tuple[float, float, float] = (0.0, 0.0, 0.0)
    radii: tuple[float, float, float] = (300.0, 200.0, 150.0)  # semi-axes

    def __post_init__(self):
        if min(self.radii) <= 0.0:
            raise ValueError("semi-axes must be positive")

    def sdf(self, x, y, z) -> np.ndarray:
        (cx, cy, cz), (a, b, c) = self.center, self.radii
        u, v, w = (x - cx) / a, (y - cy) / b, (z - cz) / c
        k1 = np.sqrt(u ** 2 + v ** 2 + w ** 2)  # implicit function + 1
        k2 = np.sqrt((u / a) ** 2 + (v / b) ** 2 + (w / c) ** 2)  # |grad| / k1 scale
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.where(k2 > 0.0, k1 * (k1 - 1.0) / k2, -min(self.radii))

    @property
    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        c = np.asarray(self.center, dtype=float)
        r = np.asarray(self.radii, dtype=float)
        return c - r, c + r

    @property
    def volume(self) -> float:
        return 4.0 / 3.0 * np.pi * float(np.prod(self.radii))


@dataclass(frozen=True)
class Union:
    shapes: tuple = field(default_factory=tuple)

    def __post_init__(self):
        if not self.shapes:
            raise ValueError("a union needs at least one shape")

    def sdf(self, x, y, z) -> np.ndarray:
        return np.minimum.reduce([s.sdf(x, y, z) for s in self.shapes])

    @property
    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        lows, highs = zip(*(s.bounds for s in self.shapes))
        return np.minimum.reduce(lows), np.maximum.reduce(highs)

    @property
    def volume(self) -> float:
        return float("nan")  # overlaps are not known analytically; count voxels instead


def evaluate_on_grid(shape, geometry: GridGeometry) -> np.ndarray:
    """The shape's SDF at every node of ``geometry``, float32 (nx, ny, nz)."""
    x, y, z = np.meshgrid(*geometry.node_coordinates(), indexing="ij")
    return shape.sdf(x, y, z).astype(np.float32)
