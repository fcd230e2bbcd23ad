"""Drop-in ``UnboundedPoissonSolverPYFFTW3D`` on the gfx950 pruned-FFT passes.

Same constructor and methods as
``sopht/numeric/eulerian_grid_ops/poisson_solver_3d/UnboundedPoissonSolverPYFFTW3D.py:9-172``.
The plan (Green's function octant and spectral slab) lives on the device; ``rhs`` is read
only, as in the reference (which copies it into its doubled buffer first).
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from ... import _lib as L


class UnboundedPoissonSolverPYFFTW3D:
    """Free-space Poisson solver ``-lap(psi) = rhs`` by domain doubling."""

    def __init__(
        self,
        grid_size_z: int,
        grid_size_y: int,
        grid_size_x: int,
        x_range: float = 1.0,
        num_threads: int = 1,
        real_t: type = np.float64,
    ) -> None:
        self.grid_size_z = grid_size_z
        self.grid_size_y = grid_size_y
        self.grid_size_x = grid_size_x
        self.x_range = x_range
        self.y_range = x_range * (grid_size_y / grid_size_x)
        self.z_range = x_range * (grid_size_z / grid_size_x)
        self.dx = real_t(x_range / grid_size_x)
        self.num_threads = num_threads
        self.real_t = real_t
        self._tag = L.dtype_tag(real_t)
        plan = C.c_void_p()
        L.call(
            "sopht_poisson3d_create", self._tag, grid_size_z, grid_size_y, grid_size_x,
            float(x_range), L.stream_ptr(), C.byref(plan),
        )
        self._plan = plan
        self._lib = L.lib()

    @property
    def plan(self):
        return self._plan

    @property
    def scratch_bytes(self) -> int:
        return int(self._lib.sopht_poisson3d_scratch_bytes(self._plan))

    @property
    def fft_schedule(self) -> dict:
        """The plan's FFT schedule: ``passes`` (3 or 5) and ``zpass_jobs`` (the three-pass z
        pass's jobs per workgroup, timed when the plan was created; 0 with five passes)."""
        passes, jobs = L.schedule("sopht_poisson3d_schedule", self._plan)
        return {"passes": passes, "zpass_jobs": jobs}

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            self._lib.sopht_poisson3d_destroy(plan)
            self._plan = None

    def _check(self, t, name):
        L.require_device_tensor(t, self.real_t, name)
        if tuple(t.shape[-3:]) != (self.grid_size_z, self.grid_size_y, self.grid_size_x):
            msg = f"{name} has shape {tuple(t.shape)}, expected (..., {self.grid_size_z}, {self.grid_size_y}, {self.grid_size_x})"
            raise ValueError(msg)

    def solve(self, solution_field, rhs_field) -> None:
        """``solve`` (``:111-149``)."""
        self._check(solution_field, "solution_field")
        self._check(rhs_field, "rhs_field")
        L.call(
            "sopht_poisson3d_solve", self._plan, L.ptr(solution_field), L.strides3(solution_field),
            L.ptr(rhs_field), L.strides3(rhs_field), L.stream_ptr(),
        )
        L.written(solution_field)

    def vector_field_solve(self, solution_vector_field, rhs_vector_field) -> None:
        """``vector_field_solve`` (``:151-172``): the three components in one pass sequence."""
        self._check(solution_vector_field, "solution_vector_field")
        self._check(rhs_vector_field, "rhs_vector_field")
        L.call(
            "sopht_poisson3d_vector_solve", self._plan, L.comps(solution_vector_field),
            L.strides3(solution_vector_field[0]), L.comps(rhs_vector_field),
            L.strides3(rhs_vector_field[0]), L.stream_ptr(),
        )
        L.written(solution_vector_field)
