"""Oracle: Eulerian-Lagrangian transfer and virtual-boundary forcing (test only).

Restates ``sopht/numeric/immersed_boundary_ops/EulerianLagrangianGridCommunicator3D.py``
and ``VirtualBoundaryForcing.py``. Arrays follow the reference shapes:
positions ``(3, n)``, nearest index ``(3, n)`` int (x, y, z rows), support
``(3, 4, 4, 4, n)`` and weights ``(4, 4, 4, n)`` with the 4^3 block indexed ``[kz, ky, kx]``.

Pinning (tests/test_oracle_ib.py): the closed forms of the reference tests
(``test_eulerian_lagrangian_grid_communicator_3d.py:7-421``,
``test_virtual_boundary_forcing.py:7-420``), partition of unity of the cosine and
Peskin kernels, and the transpose identity ``<interp(u), F> = dx^3 <u, spread(F)>``.
"""

from __future__ import annotations

import numpy as np

WIDTH = 2  # interp_kernel_width; the cosine and Peskin kernels require 2 (:397-399,425-427)


def support_offsets(width=WIDTH):
    """Integer offsets of the local support, ``arange(-w+1, w+1)`` (``:84``)."""
    return np.arange(-width + 1, width + 1)


def local_eulerian_grid_support(lag_positions, dx, eul_grid_coord_shift, real_t, width=WIDTH):
    """``:102-135``: ``idx = floor((X - shift)/dx)`` per axis; ``d[a, kz, ky, kx, n] =
    (idx_a + off_a) dx + shift - X_a`` where off_a is the offset along axis a
    (x uses kx, y uses ky, z uses kz)."""
    n = lag_positions.shape[1]
    idx = np.floor_divide(lag_positions - eul_grid_coord_shift, dx).astype(np.int64)
    off = support_offsets(width)
    k = 2 * width
    oz, oy, ox = np.meshgrid(off, off, off, indexing="ij")
    offs = np.stack((ox, oy, oz)).reshape(3, k, k, k, 1)
    d = (idx.reshape(3, 1, 1, 1, n) + offs) * dx + eul_grid_coord_shift - lag_positions.reshape(
        3, 1, 1, 1, n
    )
    return idx, d.astype(real_t)


def cosine_weights(support, dx, real_t):
    """``:401-410``: ``w = (0.25/dx)^3 prod_a (1 + cos(pi/2 * d_a/dx))`` (``d`` is divided by
    dx in place by the reference; here a copy is returned alongside)."""
    s = (support / dx).astype(real_t)
    h = real_t(0.5 * np.pi)
    w = (
        real_t((0.25 / dx) ** 3)
        * (real_t(1.0) + np.cos(h * s[0]))
        * (real_t(1.0) + np.cos(h * s[1]))
        * (real_t(1.0) + np.cos(h * s[2]))
    )
    return w.astype(real_t), s


def _peskin_1d(r):
    r = np.fabs(r)
    a = (r < 1.0) * (3.0 - 2 * r + np.sqrt(np.fabs(1 + 4 * r - 4 * r**2)))
    b = (r >= 1.0) * (r < 2.0) * (5.0 - 2 * r - np.sqrt(np.fabs(-7 + 12 * r - 4 * r**2)))
    return a + b


def peskin_weights(support, dx, real_t):
    """``:429-516``: Peskin (2002, eq. 6.27) 4-point kernel, ``(0.125/dx)^3 prod phi(|d|/dx)``."""
    s = np.fabs(support) / dx
    w = (0.125 / dx) ** 3 * _peskin_1d(s[0]) * _peskin_1d(s[1]) * _peskin_1d(s[2])
    return w.astype(real_t), s.astype(real_t)


def interpolate(eul_grid_field, interp_weights, nearest_idx, dx, width=WIDTH):
    """Eulerian -> Lagrangian (``:198-295``): ``U_n = dx^3 sum_{4^3} u[block(n)] w[..., n]``;
    the block spans ``idx - w + 1 .. idx + w`` per axis. Scalar or vector fields."""
    n = nearest_idx.shape[1]
    vec = eul_grid_field.ndim == 4
    out = np.zeros((3, n) if vec else (n,), dtype=eul_grid_field.dtype)
    for i in range(n):
        ix, iy, iz = nearest_idx[:, i]
        blk = (
            slice(iz - width + 1, iz + width + 1),
            slice(iy - width + 1, iy + width + 1),
            slice(ix - width + 1, ix + width + 1),
        )
        if vec:
            for c in range(3):
                out[c, i] = np.sum(eul_grid_field[c][blk] * interp_weights[..., i]) * dx**3
        else:
            out[i] = np.sum(eul_grid_field[blk] * interp_weights[..., i]) * dx**3
    return out


def spread(eul_grid_field, lag_grid_field, interp_weights, nearest_idx, width=WIDTH):
    """Lagrangian -> Eulerian (``:318-380``): for nodes n = 0..N-1 in order,
    ``F[block(n)] += F_lag[n] * w[..., n]`` (no dx^3). Accumulates into ``eul_grid_field``."""
    n = nearest_idx.shape[1]
    vec = eul_grid_field.ndim == 4
    for i in range(n):
        ix, iy, iz = nearest_idx[:, i]
        blk = (
            slice(iz - width + 1, iz + width + 1),
            slice(iy - width + 1, iy + width + 1),
            slice(ix - width + 1, ix + width + 1),
        )
        if vec:
            for c in range(3):
                eul_grid_field[c][blk] += lag_grid_field[c, i] * interp_weights[..., i]
        else:
            eul_grid_field[blk] += lag_grid_field[i] * interp_weights[..., i]


class VirtualBoundaryForcing:
    """Oracle of ``VirtualBoundaryForcing`` (``VirtualBoundaryForcing.py:20-283``), 3D only."""

    def __init__(
        self,
        virtual_boundary_stiffness_coeff,
        virtual_boundary_damping_coeff,
        dx,
        num_lag_nodes,
        real_t,
        eul_grid_coord_shift=None,
        interp_kernel_type="cosine",
        enable_eul_grid_forcing_reset=True,
        start_time=0.0,
    ):
        self.k = virtual_boundary_stiffness_coeff
        self.c = virtual_boundary_damping_coeff
        self.dx = dx
        self.real_t = real_t
        self.shift = real_t(dx / 2) if eul_grid_coord_shift is None else eul_grid_coord_shift
        self.kernel_type = interp_kernel_type
        self.reset = enable_eul_grid_forcing_reset
        self.time = start_time
        n = num_lag_nodes
        self.nearest_eul_grid_index_to_lag_grid = np.empty((3, n), dtype=np.int64)
        self.local_eul_grid_support_of_lag_grid = np.empty((3, 4, 4, 4, n), dtype=real_t)
        self.interp_weights = np.empty((4, 4, 4, n), dtype=real_t)
        self.lag_grid_flow_velocity_field = np.zeros((3, n), dtype=real_t)
        self.lag_grid_position_mismatch_field = np.zeros((3, n), dtype=real_t)
        self.lag_grid_velocity_mismatch_field = np.zeros((3, n), dtype=real_t)
        self.lag_grid_forcing_field = np.zeros((3, n), dtype=real_t)

    def compute_interaction_force_on_lag_grid(self, eul_grid_velocity_field, lag_pos, lag_vel):
        """``:187-230``: support, weights, interpolation, mismatch, penalty force."""
        idx, d = local_eulerian_grid_support(lag_pos, self.dx, self.shift, self.real_t)
        self.nearest_eul_grid_index_to_lag_grid[...] = idx
        wfun = cosine_weights if self.kernel_type == "cosine" else peskin_weights
        w, s = wfun(d, self.dx, self.real_t)
        self.local_eul_grid_support_of_lag_grid[...] = s
        self.interp_weights[...] = w
        self.lag_grid_flow_velocity_field[...] = interpolate(
            eul_grid_velocity_field, w, idx, self.dx
        )
        self.lag_grid_velocity_mismatch_field[...] = self.lag_grid_flow_velocity_field - lag_vel
        self.lag_grid_forcing_field[...] = (
            self.k * self.lag_grid_position_mismatch_field
            + self.c * self.lag_grid_velocity_mismatch_field
        )

    def compute_interaction_forcing(self, eul_grid_forcing_field, eul_grid_velocity_field, lag_pos, lag_vel):
        """``:232-274``: optional reset of the Eulerian forcing, then lag force and spread."""
        if self.reset:
            eul_grid_forcing_field[...] = 0
        self.compute_interaction_force_on_lag_grid(eul_grid_velocity_field, lag_pos, lag_vel)
        spread(
            eul_grid_forcing_field,
            self.lag_grid_forcing_field,
            self.interp_weights,
            self.nearest_eul_grid_index_to_lag_grid,
        )

    def time_step(self, dt):
        """``:276-283``: ``dX += dt * dV`` and ``time += dt``."""
        self.lag_grid_position_mismatch_field[...] = (
            self.lag_grid_position_mismatch_field + dt * self.lag_grid_velocity_mismatch_field
        )
        self.time += dt
