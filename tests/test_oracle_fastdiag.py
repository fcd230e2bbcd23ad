"""Pins the oracle's Neumann fast-diagonalisation Poisson solve (CPU only).

* The reference's closed-form tests (``test_unbounded_poisson_solver_3d.py:160-232``):
  ``cos(4 pi x) cos(4 pi y) cos(4 pi z)`` on a 16^3 unit cube solves to itself over
  ``3 (4 pi)^2`` to 1e-4, scalar and vector, single and double.
* The discrete identity: ``-L_h psi = rhs - mean(rhs)`` and ``mean(psi) = 0``, with ``L_h``
  an independent 7-point Laplacian whose ghost cells mirror the edge cells (the Neumann
  condition at the cell faces), on anisotropic grids. A transposed basis, a wrong sign,
  a dropped axis or a null mode left in all fail it.
* The eigenbasis: the cell-centred cosines ``c_k cos(pi k (i + 1/2) / N)`` with eigenvalues
  ``(4/dx^2) sin^2(pi k / 2N)`` in decreasing order equal ``numpy.linalg.eig``'s columns up
  to sign (what the device library builds in closed form).
"""

import numpy as np
import pytest

from oracle import fastdiag as FD


def neumann_laplacian(f, dx):
    """7-point Laplacian with mirrored ghost cells (zero normal derivative at the faces)."""
    p = np.pad(f, 1, mode="edge")
    c = p[1:-1, 1:-1, 1:-1]
    lap = (
        p[2:, 1:-1, 1:-1] + p[:-2, 1:-1, 1:-1]
        + p[1:-1, 2:, 1:-1] + p[1:-1, :-2, 1:-1]
        + p[1:-1, 1:-1, 2:] + p[1:-1, 1:-1, :-2]
        - 6 * c
    )
    return lap / dx**2


def cos_field(n, real_t, wave_number=4):
    dx = real_t(1.0 / n)
    x = np.linspace(dx / 2, 1 - dx / 2, n).astype(real_t)
    z, y, xx = np.meshgrid(x, x, x, indexing="ij")
    k = real_t(wave_number)
    f = np.cos(k * np.pi * xx) * np.cos(k * np.pi * y) * np.cos(k * np.pi * z)
    return f.astype(real_t), dx, f / (3 * (k**2) * np.pi**2)


@pytest.mark.parametrize("real_t", [np.float32, np.float64])
def test_closed_form_scalar(real_t):
    rhs, dx, ref = cos_field(16, real_t)
    solver = FD.FastDiagPoissonSolver3D(16, 16, 16, dx, real_t)
    sol = np.zeros_like(rhs)
    solver.solve(sol, rhs)
    assert np.amax(np.fabs(sol - ref)) <= real_t(1e-4)


@pytest.mark.parametrize("real_t", [np.float32, np.float64])
def test_closed_form_vector(real_t):
    rhs, dx, ref = cos_field(16, real_t)
    rhs3 = np.stack([rhs] * 3)
    solver = FD.FastDiagPoissonSolver3D(16, 16, 16, dx, real_t)
    sol = np.zeros_like(rhs3)
    solver.vector_field_solve(sol, rhs3)
    assert np.amax(np.fabs(sol - np.stack([ref] * 3))) <= real_t(1e-4)


@pytest.mark.parametrize("shape", [(6, 10, 14), (12, 8, 5), (16, 16, 16)])
def test_discrete_identity(shape):
    rng = np.random.default_rng(31)
    rhs = rng.standard_normal(shape)
    dx = 0.07
    solver = FD.FastDiagPoissonSolver3D(*shape, dx)
    sol = np.zeros(shape)
    solver.solve(sol, rhs)
    res = -neumann_laplacian(sol, dx) - (rhs - rhs.mean())
    assert np.abs(res).max() <= 1e-10 * np.abs(rhs).max()
    assert abs(sol.mean()) <= 1e-12 * np.abs(sol).max()


def test_discrete_identity_fails_with_a_transposed_basis():
    """A planted mistake (the forward y transform with V instead of V^-1) breaks the identity."""
    shape = (6, 10, 14)
    rng = np.random.default_rng(32)
    rhs = rng.standard_normal(shape)
    dx = 0.07
    solver = FD.FastDiagPoissonSolver3D(*shape, dx)
    solver.inv_of_eig_vecs_y = solver.eig_vecs_y.copy()
    sol = np.zeros(shape)
    solver.solve(sol, rhs)
    res = -neumann_laplacian(sol, dx) - (rhs - rhs.mean())
    assert np.abs(res).max() > 1e-3 * np.abs(rhs).max()


@pytest.mark.parametrize("n", [2, 7, 16, 33])
def test_cosine_eigenbasis_matches_eig(n):
    dx = 0.3
    vals, vecs, _ = FD.sorted_eig(FD.neumann_poisson_matrix(n, dx, np.float64))
    k = n - 1 - np.arange(n)  # decreasing eigenvalues
    lam = 4 / dx**2 * np.sin(np.pi * k / (2 * n)) ** 2
    c = np.where(k == 0, np.sqrt(1 / n), np.sqrt(2 / n))
    i = np.arange(n).reshape(-1, 1)
    basis = c * np.cos(np.pi * k * (i + 0.5) / n)
    np.testing.assert_allclose(vals, lam, rtol=0, atol=1e-12 / dx**2)
    sign = np.sign(np.sum(basis * vecs, axis=0))
    np.testing.assert_allclose(vecs * sign, basis, rtol=0, atol=1e-12)


def test_vector_solve_is_componentwise():
    rng = np.random.default_rng(33)
    rhs = rng.standard_normal((3, 6, 8, 10))
    solver = FD.FastDiagPoissonSolver3D(6, 8, 10, 0.1)
    out = np.zeros_like(rhs)
    solver.vector_field_solve(out, rhs)
    for c in range(3):
        one = np.zeros(rhs.shape[1:])
        solver.solve(one, rhs[c])
        np.testing.assert_array_equal(one, out[c])
