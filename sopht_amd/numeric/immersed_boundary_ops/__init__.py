"""Immersed boundary operations (3D), same names as ``sopht/numeric/immersed_boundary_ops``."""

from .eulerian_lagrangian_grid_communicator_3d import EulerianLagrangianGridCommunicator3D
from .virtual_boundary_forcing import VirtualBoundaryForcing

__all__ = ["EulerianLagrangianGridCommunicator3D", "VirtualBoundaryForcing"]
