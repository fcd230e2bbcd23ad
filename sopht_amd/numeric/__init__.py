"""Numeric kernels of the 3D flow step (Eulerian grid ops and immersed-boundary ops)."""

from . import eulerian_grid_ops, immersed_boundary_ops  # noqa: F401
