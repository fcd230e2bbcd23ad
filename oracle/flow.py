"""Oracle: the 3D unbounded Navier-Stokes flow step with forcing (test only).

Restates ``UnboundedNavierStokesFlowSimulator3D`` (``sopht/simulator/flow/
navier_stokes_flow_simulators.py:226-522``), its base ``FlowSimulator``
(``flow_simulators.py:9-133``), ``compute_advection_diffusion_stable_timestep``
(``passive_transport_flow_simulators.py:139-155``) and the host glue of
``ImmersedBodyFlowInteraction`` (``immersed_body_flow_interaction.py:15-141``).

Pinning (tests/test_oracle_flow.py): the composite step equals the manual kernel sequence
of ``tests/test_simulator/test_flow/test_navier_stokes_flow_simulators.py:267-308`` built
from the separately pinned pieces; ``dt`` matches the closed form of ``:311-338``.
"""

from __future__ import annotations

import numpy as np

from . import fastdiag, ib, poisson, stencils


def position_field(grid_size, x_range, real_t):
    """``FlowSimulator._init_domain`` (``flow_simulators.py:50-95``): cell-centred
    coordinates, ``(3, Nz, Ny, Nx)`` with components x, y, z."""
    nz, ny, nx = grid_size
    dx = real_t(x_range / nx)
    shift = dx / 2.0
    y_range = x_range * ny / nx
    z_range = x_range * nz / nx
    x = np.linspace(shift, x_range - shift, nx).astype(real_t)
    y = np.linspace(shift, y_range - shift, ny).astype(real_t)
    z = np.linspace(shift, z_range - shift, nz).astype(real_t)
    return np.flipud(np.array(np.meshgrid(z, y, x, indexing="ij"))), dx


def compute_advection_diffusion_stable_timestep(velocity_field, dx, cfl, nu, real_t, grid_dim=3):
    """``passive_transport_flow_simulators.py:139-155``: ``tol = 10 eps``;
    ``min(cfl dx/(max sum|u_i| + tol), 0.9 dx^2/(2 d)/nu + tol)`` in ``real_t``."""
    tol = 10 * np.finfo(real_t).eps
    mag = np.sum(np.fabs(velocity_field), axis=0).astype(real_t)
    return min(cfl * dx / (np.amax(mag) + tol), 0.9 * dx**2 / (2 * grid_dim) / nu + tol)


class UnboundedNavierStokesFlowSimulator3D:
    """Oracle of the reference simulator (greens-function or fast-diagonalisation Poisson)."""

    def __init__(
        self,
        grid_size,
        x_range,
        kinematic_viscosity,
        cfl=0.1,
        real_t=np.float64,
        time=0.0,
        with_forcing=False,
        with_free_stream_flow=False,
        flow_density=1.0,
        penalty_zone_width=2,
        filter_setting_dict=None,
        poisson_solver_type="greens_function_convolution",
    ):
        self.grid_size = tuple(grid_size)
        self.x_range = x_range
        self.real_t = real_t
        self.kinematic_viscosity = kinematic_viscosity
        self.cfl = cfl
        self.time = time
        self.with_forcing = with_forcing
        self.with_free_stream_flow = with_free_stream_flow
        self.flow_density = flow_density
        self.penalty_zone_width = penalty_zone_width
        # navier_stokes_flow_simulators.py:274-292,396-408 (None: filter off)
        self.filter_setting_dict = filter_setting_dict
        self.position_field, self.dx = position_field(self.grid_size, x_range, real_t)
        shape = (3, *self.grid_size)
        self.vorticity_field = np.zeros(shape, dtype=real_t)
        self.velocity_field = np.zeros(shape, dtype=real_t)
        self.buffer_vector_field = np.zeros(shape, dtype=real_t)
        self.buffer_scalar_field = self.buffer_vector_field[0].view()
        self.stream_func_field = np.zeros(shape, dtype=real_t)
        if with_forcing:
            self.eul_grid_forcing_field = np.zeros(shape, dtype=real_t)
        # navier_stokes_flow_simulators.py:337-358
        if poisson_solver_type == "fast_diagonalisation":
            self.poisson = fastdiag.FastDiagPoissonSolver3D(*self.grid_size, dx=self.dx, real_t=real_t)
        else:
            self.poisson = poisson.UnboundedPoissonSolver3D(*self.grid_size, x_range=x_range, real_t=real_t)

    # -- pieces of ``_navier_stokes_time_step`` (:449-485)
    def penalise(self, field):
        pf = self.position_field
        stencils.penalise_vector_field_boundary(
            field, self.penalty_zone_width, self.dx, pf[0], pf[1], pf[2]
        )

    def recover_velocity(self, free_stream_velocity=(0.0, 0.0, 0.0)):
        """Penalise, Poisson, curl and free stream (``:475-485``)."""
        self.penalise(self.vorticity_field)
        self.poisson.vector_field_solve(self.stream_func_field, self.vorticity_field)
        stencils.curl(self.velocity_field, self.stream_func_field, self.real_t(0.5 / self.dx))
        if self.with_free_stream_flow:
            stencils.add_fixed_val(self.velocity_field, self.velocity_field, np.asarray(free_stream_velocity))

    def navier_stokes_time_step(self, dt, free_stream_velocity=(0.0, 0.0, 0.0)):
        """``_navier_stokes_time_step`` (``:449-485``); the vorticity filter (``:473``) when
        ``filter_setting_dict`` is given."""
        rt = self.real_t
        b = self.buffer_vector_field
        stencils.elementwise_cross_product(b, self.velocity_field, self.vorticity_field)
        stencils.update_vorticity_from_velocity_forcing(self.vorticity_field, b, rt(dt / (2 * self.dx)))
        stencils.diffusion_timestep_euler_forward(
            self.vorticity_field,
            self.buffer_scalar_field,
            rt(self.kinematic_viscosity * dt / self.dx / self.dx),
        )
        if self.filter_setting_dict is not None:  # :473
            stencils.laplacian_filter_vector(
                self.vorticity_field, self.filter_setting_dict["order"], self.filter_setting_dict["type"],
                self.buffer_vector_field[0], self.buffer_vector_field[1],
            )
        self.recover_velocity(free_stream_velocity)

    def time_step(self, dt, free_stream_velocity=(0.0, 0.0, 0.0)):
        """``FlowSimulator.time_step`` (``flow_simulators.py:130-133``) bound to
        ``_navier_stokes_with_forcing_time_step`` (``:487-498``) when forcing is on."""
        if self.with_forcing:
            stencils.update_vorticity_from_velocity_forcing(
                self.vorticity_field,
                self.eul_grid_forcing_field,
                self.real_t(dt / (2 * self.dx * self.flow_density)),
            )
        self.navier_stokes_time_step(dt, free_stream_velocity)
        if self.with_forcing:
            self.eul_grid_forcing_field[...] = 0
        self.time += dt

    def compute_stable_timestep(self, dt_prefac=1.0):
        """``:501-512``."""
        return dt_prefac * compute_advection_diffusion_stable_timestep(
            self.velocity_field, self.dx, self.cfl, self.kinematic_viscosity, self.real_t
        )

    def get_vorticity_divergence_l2_norm(self):
        """``:514-522``."""
        div = np.zeros(self.grid_size, dtype=self.real_t)
        stencils.divergence(div, self.vorticity_field, 1.0 / self.dx)
        return np.linalg.norm(div) * self.dx ** (3 / 2.0)


class ImmersedBodyFlowInteraction(ib.VirtualBoundaryForcing):
    """Host glue of ``ImmersedBodyFlowInteraction`` (``immersed_body_flow_interaction.py:15-141``):
    coefficients rescaled by ``max_lag_grid_dx**(d-1)`` (``:83-86``); forcing reset off by
    default (``:33``); ``__call__`` (``:113-122``)."""

    def __init__(self, flow, forcing_grid, k, c, real_t, enable_eul_grid_forcing_reset=False,
                 body_flow_forces=None, body_flow_torques=None):
        self.body_flow_forces = np.zeros((3, 1)) if body_flow_forces is None else body_flow_forces
        self.body_flow_torques = np.zeros((3, 1)) if body_flow_torques is None else body_flow_torques
        ds = forcing_grid.get_maximum_lagrangian_grid_spacing()
        super().__init__(
            k * ds**2,
            c * ds**2,
            flow.dx,
            forcing_grid.num_lag_nodes,
            real_t,
            enable_eul_grid_forcing_reset=enable_eul_grid_forcing_reset,
        )
        self.flow = flow
        self.forcing_grid = forcing_grid

    def __call__(self):
        self.forcing_grid.compute_lag_grid_position_field()
        self.forcing_grid.compute_lag_grid_velocity_field()
        self.compute_interaction_forcing(
            self.flow.eul_grid_forcing_field,
            self.flow.velocity_field,
            self.forcing_grid.position_field,
            self.forcing_grid.velocity_field,
        )

    def compute_interaction_on_lag_grid(self):
        """``:102-111``: grid kinematics, then the Lagrangian forcing alone (no spreading)."""
        self.forcing_grid.compute_lag_grid_position_field()
        self.forcing_grid.compute_lag_grid_velocity_field()
        self.compute_interaction_force_on_lag_grid(
            self.flow.velocity_field, self.forcing_grid.position_field, self.forcing_grid.velocity_field
        )

    def compute_flow_forces_and_torques(self):
        """``:124-131``: the Lagrangian forcing, transferred to the body by its grid."""
        self.compute_interaction_on_lag_grid()
        self.forcing_grid.transfer_forcing_from_grid_to_body(
            self.body_flow_forces, self.body_flow_torques, self.lag_grid_forcing_field
        )
