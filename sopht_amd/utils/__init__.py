"""Config plumbing shared by the numeric and simulator layers."""

from .field import VectorField
from .precision import get_real_t, get_test_tol

__all__ = ["VectorField", "get_real_t", "get_test_tol"]
