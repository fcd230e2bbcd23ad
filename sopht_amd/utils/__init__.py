"""Config plumbing shared by the numeric and simulator layers."""

from .field import VectorField
from .io import IO, EulerianFieldIO
from .precision import get_real_t, get_test_tol
from .restart_sim import restart_simulation

__all__ = ["IO", "EulerianFieldIO", "VectorField", "get_real_t", "get_test_tol", "restart_simulation"]
