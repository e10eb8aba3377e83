"""Tissue regions: synthetic shapes as signed-distance fields and the region the contact pass
confines cells to (geometry/ of the reference: shapes.py, region.py)."""

from warpbiocell_amd.geometry.region import TissueRegion, node_gradient
from warpbiocell_amd.geometry.shapes import Ellipsoid, Sphere, Union, evaluate_on_grid

__all__ = ["TissueRegion", "node_gradient", "Sphere", "Ellipsoid", "Union", "evaluate_on_grid"]
