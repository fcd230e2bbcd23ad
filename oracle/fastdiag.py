"""Oracle: Neumann Poisson solve by fast diagonalisation (test only).

Restates ``sopht/numeric/eulerian_grid_ops/poisson_solver_3d/FastDiagPoissonSolver3D.py``:
per axis the second-difference matrix ``(1/dx^2) tridiag(-1, 2, -1)`` with its corner
entries set to ``1/dx^2`` (``:53-96``), a dense eigendecomposition (``numpy.linalg.eig``)
with the eigenvalues sorted in decreasing order and the inverse eigenvector matrix from
``numpy.linalg.inv`` (``:98-125``), ``1 / (lz + ly + lx)`` with the last (null) mode set to
0 (``:127-145``), and the solve as ``tensordot`` along x, y, z, the spectral product, and
``tensordot`` back along x, y, z (``:147-189``).

Pinning (tests/test_oracle_fastdiag.py): the reference's closed-form tests
(``test_unbounded_poisson_solver_3d.py:160-232``: ``cos(4 pi x) cos(4 pi y) cos(4 pi z)``
maps to itself over ``3 (4 pi)^2``, to 1e-4); the discrete identity
``-L_h psi = rhs - mean(rhs)``, ``mean(psi) = 0`` with ``L_h`` an independent 7-point
Neumann Laplacian (mirrored ghost cells) on anisotropic grids; and the closed-form cosine
eigenbasis of the Neumann matrix against ``numpy.linalg.eig`` up to column signs.
"""

from __future__ import annotations

import numpy as np
import numpy.linalg as la


def neumann_poisson_matrix(n, dx, real_t):
    """``_construct_poisson_matrices`` and ``_apply_boundary_conds_to_poisson_matrices``
    (``:53-96``) for one axis."""
    inv_dx2 = real_t(1 / dx / dx)
    m = np.zeros((n, n), dtype=real_t)
    idx = np.arange(n)
    m[idx, idx] = 2 * inv_dx2
    m[idx[:-1], idx[:-1] + 1] = -inv_dx2
    m[idx[1:], idx[1:] - 1] = -inv_dx2
    m[0, 0] = inv_dx2
    m[-1, -1] = inv_dx2
    return m


def sorted_eig(m):
    """``:98-125``: eigenpairs sorted by decreasing eigenvalue, and the inverse of the
    eigenvector matrix."""
    vals, vecs = la.eig(m)
    order = vals.argsort()[::-1]
    vals = vals[order]
    vecs = vecs[:, order]
    return vals, vecs, la.inv(vecs)


class FastDiagPoissonSolver3D:
    """Oracle of ``FastDiagPoissonSolver3D`` (``:12-202``), bc_type
    ``homogenous_neumann_along_xyz``."""

    def __init__(self, grid_size_z, grid_size_y, grid_size_x, dx, real_t=np.float64):
        self.grid_size = (grid_size_z, grid_size_y, grid_size_x)
        self.dx = dx
        self.real_t = real_t
        vz, self.eig_vecs_z, self.inv_of_eig_vecs_z = sorted_eig(neumann_poisson_matrix(grid_size_z, dx, real_t))
        vy, self.eig_vecs_y, self.inv_of_eig_vecs_y = sorted_eig(neumann_poisson_matrix(grid_size_y, dx, real_t))
        vx, self.eig_vecs_x, self.inv_of_eig_vecs_x = sorted_eig(neumann_poisson_matrix(grid_size_x, dx, real_t))
        self.eig_vals = (vz, vy, vx)
        lam = vz.reshape(-1, 1, 1) + vy.reshape(1, -1, 1) + vx.reshape(1, 1, -1)
        lam[-1, -1, -1] = np.inf
        self.inv_eig_val_matrix = real_t(1) / lam

    def solve(self, solution_field, rhs_field):
        """``solve`` (``:147-189``)."""
        s = np.tensordot(rhs_field, self.inv_of_eig_vecs_x, axes=(2, 1))
        s = np.tensordot(self.inv_of_eig_vecs_y, s, axes=(1, 1)).transpose((1, 0, 2))
        s = np.tensordot(self.inv_of_eig_vecs_z, s, axes=(1, 0))
        s = s * self.inv_eig_val_matrix
        s = np.tensordot(s, self.eig_vecs_x, axes=(2, 1))
        s = np.tensordot(self.eig_vecs_y, s, axes=(1, 1)).transpose((1, 0, 2))
        solution_field[...] = np.tensordot(self.eig_vecs_z, s, axes=(1, 0))

    def vector_field_solve(self, solution_vector_field, rhs_vector_field):
        """``vector_field_solve`` (``:191-202``): x, y, z components in turn."""
        for c in range(3):
            self.solve(solution_vector_field[c], rhs_vector_field[c])
