"""Oracle: free-space Poisson solve ``-lap(psi) = omega`` by domain doubling (test only).

Restates ``sopht/numeric/eulerian_grid_ops/poisson_solver_3d/UnboundedPoissonSolverPYFFTW3D.py``
with ``scipy.fft.rfftn/irfftn`` in place of FFTW (the backward transform normalised by
``1/(2N)^3``, pyfftw's default, pinned by ``test_fft_kernel_3d.py:57-67``).

Pinning (tests/test_oracle_poisson.py): a brute-force free-space convolution
``psi[n] = dx^3 sum_m G(|n-m|) omega[m]`` on tiny grids (which fixes the doubling,
the crop, the Green's function, its regularisation and the normalisation), and
convergence to the analytic potential of a Gaussian vorticity blob.
"""

from __future__ import annotations

import numpy as np
from scipy.fft import irfftn, rfftn

# scipy.fft worker threads (the CPU baseline sets this to the host's core count)
WORKERS = 1


def fourier_greens_function(grid_size, x_range, real_t):
    """``_construct_fourier_greens_function_field`` (``:56-83``) times ``dx^3`` (``:47-49``).

    G = 1/(4 pi r_even) on the doubled grid ``linspace(0, 2L - dx, 2N)`` per axis with
    ``r_even`` from ``min(x, 2L - x)``; G(0) = 1/(4 pi dx). Computed in ``real_t`` as the
    reference does; the result is the rfftn of G times dx^3 (complex).
    """
    nz, ny, nx = grid_size
    dx = real_t(x_range / nx)
    y_range = x_range * (ny / nx)
    z_range = x_range * (nz / nx)
    xd = np.linspace(0, 2 * x_range - dx, 2 * nx).astype(real_t)
    yd = np.linspace(0, 2 * y_range - dx, 2 * ny).astype(real_t)
    zd = np.linspace(0, 2 * z_range - dx, 2 * nz).astype(real_t)
    zg, yg, xg = np.meshgrid(zd, yd, xd, indexing="ij")
    r = np.sqrt(
        np.minimum(xg, 2 * x_range - xg) ** 2
        + np.minimum(yg, 2 * y_range - yg) ** 2
        + np.minimum(zg, 2 * z_range - zg) ** 2
    )
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        g = (1 / r) / (4 * np.pi)
    g[0, 0, 0] = 1 / (4 * np.pi * dx)
    return rfftn(g, workers=WORKERS) * (dx**3)


class UnboundedPoissonSolver3D:
    """Oracle of ``UnboundedPoissonSolverPYFFTW3D`` (``:9-172``)."""

    def __init__(self, grid_size_z, grid_size_y, grid_size_x, x_range=1.0, real_t=np.float64):
        self.grid_size = (grid_size_z, grid_size_y, grid_size_x)
        self.real_t = real_t
        self.dx = real_t(x_range / grid_size_x)
        self.ghat_dx3 = fourier_greens_function(self.grid_size, x_range, real_t)

    def solve(self, solution_field, rhs_field):
        """``solve`` (``:111-149``): zero the doubled buffer, copy rhs into the low octant,
        rfft, multiply by G^ dx^3, irfft, crop."""
        nz, ny, nx = self.grid_size
        buf = np.zeros((2 * nz, 2 * ny, 2 * nx), dtype=self.real_t)
        buf[:nz, :ny, :nx] = rhs_field
        out = irfftn(rfftn(buf, workers=WORKERS) * self.ghat_dx3, s=buf.shape, workers=WORKERS)
        solution_field[...] = out[:nz, :ny, :nx]

    def vector_field_solve(self, solution_vector_field, rhs_vector_field):
        """``vector_field_solve`` (``:151-172``): components x, y, z in turn."""
        for i in range(3):
            self.solve(solution_vector_field[i], rhs_vector_field[i])


def brute_force_free_space_solve(rhs_field, dx):
    """Direct O(N^6) free-space convolution ``psi[n] = dx^3 sum_m G(n - m) rhs[m]`` with
    ``G(d) = 1/(4 pi |d| dx)`` and ``G(0) = 1/(4 pi dx)``. Independent of any FFT; used
    to pin the oracle (and thereby the GPU path) on tiny grids."""
    nz, ny, nx = rhs_field.shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pts = np.stack([z.ravel(), y.ravel(), x.ravel()], axis=1)
    diff = pts[:, None, :] - pts[None, :, :]
    r = np.sqrt((diff.astype(np.float64) ** 2).sum(-1)) * dx
    with np.errstate(divide="ignore"):
        g = 1.0 / (4 * np.pi * r)
    np.fill_diagonal(g, 1.0 / (4 * np.pi * dx))
    return (g @ rhs_field.ravel().astype(np.float64) * dx**3).reshape(rhs_field.shape)
