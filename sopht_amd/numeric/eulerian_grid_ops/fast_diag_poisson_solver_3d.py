"""Drop-in ``FastDiagPoissonSolver3D`` on rocBLAS GEMMs (gfx950 MFMA).

Same constructor and methods as
``sopht/numeric/eulerian_grid_ops/poisson_solver_3d/FastDiagPoissonSolver3D.py:12-202``:
``-lap(solution) = rhs`` with homogeneous Neumann conditions on every face, the constant
mode removed. The per-axis eigenbasis of the Neumann second-difference matrix is the
cell-centred cosine basis, built in closed form in double precision by the library
(orthonormal, eigenvalues in decreasing order as in the reference's ``:98-145``); the six
per-axis contractions and the spectral scaling run on the device.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from ... import _lib as L

_AXES = {"z": 0, "y": 1, "x": 2}


class FastDiagPoissonSolver3D:
    """Neumann Poisson solver in 3D by fast diagonalisation."""

    def __init__(
        self,
        grid_size_z: int,
        grid_size_y: int,
        grid_size_x: int,
        dx: float,
        real_t: type = np.float64,
        bc_type: str = "homogenous_neumann_along_xyz",
    ) -> None:
        if bc_type != "homogenous_neumann_along_xyz":
            msg = f"Unsupported bc_type {bc_type!r}; only 'homogenous_neumann_along_xyz' is defined"
            raise ValueError(msg)
        self.dx = dx
        self.grid_size_z = grid_size_z
        self.grid_size_y = grid_size_y
        self.grid_size_x = grid_size_x
        self.real_t = real_t
        self.bc_type = bc_type
        self._tag = L.dtype_tag(real_t)
        plan = C.c_void_p()
        L.call(
            "sopht_fastdiag3d_create", self._tag, grid_size_z, grid_size_y, grid_size_x,
            float(dx), L.stream_ptr(), C.byref(plan),
        )
        self._plan = plan
        self._lib = L.lib()
        # eigenbasis on the host, as the reference's attributes (:108-125)
        for name, axis in _AXES.items():
            n = (grid_size_z, grid_size_y, grid_size_x)[axis]
            vals = np.zeros(n, dtype=np.float64)
            vecs = np.zeros((n, n), dtype=np.float64)
            L.call(
                "sopht_fastdiag3d_eigenbasis", plan, axis,
                vals.ctypes.data_as(C.POINTER(C.c_double)),
                vecs.ctypes.data_as(C.POINTER(C.c_double)),
            )
            setattr(self, f"eig_vals_{name}", vals.astype(real_t))
            setattr(self, f"eig_vecs_{name}", vecs.astype(real_t))
            # orthonormal basis: the inverse is the transpose
            setattr(self, f"inv_of_eig_vecs_{name}", np.ascontiguousarray(vecs.T).astype(real_t))
        self._inv_eig_val_matrix = None

    @property
    def inv_eig_val_matrix(self):
        """``1 / (lz + ly + lx)`` with the null mode at ``[-1, -1, -1]`` set to 0 (``:127-145``),
        a device tensor built on first access."""
        if self._inv_eig_val_matrix is None:
            import torch

            rt = self.real_t
            lam = (
                self.eig_vals_z.astype(rt).reshape(-1, 1, 1)
                + self.eig_vals_y.astype(rt).reshape(1, -1, 1)
                + self.eig_vals_x.astype(rt).reshape(1, 1, -1)
            )
            lam[-1, -1, -1] = np.inf
            inv = (rt(1) / lam).astype(rt)
            self._inv_eig_val_matrix = torch.from_numpy(inv).to(
                torch.device("cuda", torch.cuda.current_device())
            )
        return self._inv_eig_val_matrix

    @property
    def plan(self):
        return self._plan

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            self._lib.sopht_fastdiag3d_destroy(plan)
            self._plan = None

    def _check(self, t, name):
        L.require_device_tensor(t, self.real_t, name)
        if tuple(t.shape[-3:]) != (self.grid_size_z, self.grid_size_y, self.grid_size_x):
            msg = f"{name} has shape {tuple(t.shape)}, expected (..., {self.grid_size_z}, {self.grid_size_y}, {self.grid_size_x})"
            raise ValueError(msg)

    def solve(self, solution_field, rhs_field) -> None:
        """``solve`` (``:147-189``). Non-contiguous views go through contiguous copies."""
        self._check(solution_field, "solution_field")
        self._check(rhs_field, "rhs_field")
        rhs = rhs_field if rhs_field.is_contiguous() else rhs_field.contiguous()
        sol = solution_field if solution_field.is_contiguous() else solution_field.contiguous()
        L.call(
            "sopht_fastdiag3d_solve", self._plan, L.ptr(sol), L.strides3(sol), L.ptr(rhs),
            L.strides3(rhs), L.stream_ptr(),
        )
        if sol is not solution_field:
            solution_field.copy_(sol)
        else:
            L.written(solution_field)

    def vector_field_solve(self, solution_vector_field, rhs_vector_field) -> None:
        """``vector_field_solve`` (``:191-202``): the three components in one batched sequence
        when the fields are contiguous ``(3, nz, ny, nx)`` tensors."""
        self._check(solution_vector_field, "solution_vector_field")
        self._check(rhs_vector_field, "rhs_vector_field")
        if solution_vector_field.shape[0] != 3 or rhs_vector_field.shape[0] != 3:
            msg = "vector fields must have 3 components"
            raise ValueError(msg)
        rhs = rhs_vector_field if rhs_vector_field.is_contiguous() else rhs_vector_field.contiguous()
        sol = (
            solution_vector_field
            if solution_vector_field.is_contiguous()
            else solution_vector_field.contiguous()
        )
        L.call(
            "sopht_fastdiag3d_vector_solve", self._plan, L.comps(sol), L.strides3(sol[0]),
            L.comps(rhs), L.strides3(rhs[0]), L.stream_ptr(),
        )
        if sol is not solution_vector_field:
            solution_vector_field.copy_(sol)
        else:
            L.written(solution_vector_field)
