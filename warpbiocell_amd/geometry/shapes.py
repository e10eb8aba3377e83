"""Analytic tissue shapes for ``TissueRegion``: signed distances on grid nodes, negative inside.

The reference's shape API (geometry/shapes.py:17-92) is kept: ``Sphere(center, radius)``,
``Ellipsoid(center, radii)`` and ``Union(shapes)``, each with ``sdf(x, y, z)``, ``bounds`` and
``volume``, plus ``evaluate_on_grid``. Here every shape is built on one operation,
``signed_distance(points)`` over an array of points (..., 3) in um; ``sdf`` stacks its coordinate
arrays into such an array. The distance formulas are the ones the reference's regions are made
from (a region's confinement and pinned nodes depend on them):

- sphere: the exact distance |p - c| - r;
- ellipsoid: with q = (p - c) / radii, the estimate |q| (|q| - 1) / |q / radii|, the implicit
  function |q| - 1 divided by its gradient's length to first order; exact on the axes and at the
  surface, and taken as -min(radii) at the centre, where the gradient vanishes;
- union: the smallest member distance.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from warpbiocell_amd.fields.scalar_field import GridGeometry


def _as_points(x, y, z) -> np.ndarray:
    return np.stack(np.broadcast_arrays(np.asarray(x, float), np.asarray(y, float), np.asarray(z, float)), axis=-1)


class _Shape:
    """Shared part of the shapes: ``sdf`` on coordinate arrays from ``signed_distance`` on points."""

    def signed_distance(self, points: np.ndarray) -> np.ndarray:  # pragma: no cover - abstract
        raise NotImplementedError

    def sdf(self, x, y, z) -> np.ndarray:
        """Signed distance at the points (x, y, z) (arrays of one shape, or broadcastable)."""
        return self.signed_distance(_as_points(x, y, z))

    def contains(self, points: np.ndarray) -> np.ndarray:
        return self.signed_distance(np.asarray(points, float)) < 0.0


@dataclass(frozen=True)
class Sphere(_Shape):
    center: tuple[float, float, float] = (0.0, 0.0, 0.0)
    radius: float = 200.0

    def __post_init__(self):
        if not self.radius > 0.0:
            raise ValueError(f"a sphere needs a positive radius, got {self.radius}")

    def signed_distance(self, points: np.ndarray) -> np.ndarray:
        offset = np.asarray(points, float) - np.asarray(self.center, float)
        return np.sqrt(np.sum(offset * offset, axis=-1)) - self.radius

    @property
    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        c = np.asarray(self.center, float)
        return c - self.radius, c + self.radius

    @property
    def volume(self) -> float:
        return float(4.0 / 3.0 * np.pi * self.radius ** 3)


@dataclass(frozen=True)
class Ellipsoid(_Shape):
    center: tuple[float, float, float] = (0.0, 0.0, 0.0)
    radii: tuple[float, float, float] = (300.0, 200.0, 150.0)  # semi-axes along x, y, z

    def __post_init__(self):
        if len(self.radii) != 3 or not min(self.radii) > 0.0:
            raise ValueError(f"an ellipsoid needs three positive semi-axes, got {self.radii}")

    def signed_distance(self, points: np.ndarray) -> np.ndarray:
        semi = np.asarray(self.radii, float)
        q = (np.asarray(points, float) - np.asarray(self.center, float)) / semi
        level = np.sqrt(np.sum(q * q, axis=-1))        # |q|: 1 on the surface
        slope = np.sqrt(np.sum((q / semi) ** 2, axis=-1))  # |grad |q|| times |q|
        out = np.full(level.shape, -float(semi.min()))
        np.divide(level * (level - 1.0), slope, out=out, where=slope > 0.0)
        return out

    @property
    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        c, semi = np.asarray(self.center, float), np.asarray(self.radii, float)
        return c - semi, c + semi

    @property
    def volume(self) -> float:
        a, b, c = self.radii
        return float(4.0 / 3.0 * np.pi * a * b * c)


@dataclass(frozen=True)
class Union(_Shape):
    shapes: tuple = field(default_factory=tuple)

    def __post_init__(self):
        if len(self.shapes) == 0:
            raise ValueError("a union of no shapes is empty")

    def signed_distance(self, points: np.ndarray) -> np.ndarray:
        points = np.asarray(points, float)
        out = self.shapes[0].signed_distance(points)
        for s in self.shapes[1:]:
            out = np.minimum(out, s.signed_distance(points))
        return out

    @property
    def bounds(self) -> tuple[np.ndarray, np.ndarray]:
        lo = np.min([s.bounds[0] for s in self.shapes], axis=0)
        hi = np.max([s.bounds[1] for s in self.shapes], axis=0)
        return lo, hi

    @property
    def volume(self) -> float:
        """Not defined in closed form: members may overlap (NaN, as the reference returns); a
        region's volume counts its inside voxels."""
        return float("nan")


def evaluate_on_grid(shape, geometry: GridGeometry) -> np.ndarray:
    """The shape's signed distance at every node of ``geometry`` as float32 (nx, ny, nz)."""
    x, y, z = geometry.node_coordinates()
    points = _as_points(x[:, None, None], y[None, :, None], z[None, None, :])
    return shape.signed_distance(points).astype(np.float32)
