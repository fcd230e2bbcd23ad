"""Axis indices of a vector field ``(3, Nz, Ny, Nx)`` (``sopht/utils/field.py:1-39``)."""


class VectorField:
    """Component index of each axis in a vector field: x = 0, y = 1, z = 2."""

    @staticmethod
    def x_axis_idx() -> int:
        return 0

    @staticmethod
    def y_axis_idx() -> int:
        return 1

    @staticmethod
    def z_axis_idx() -> int:
        return 2
