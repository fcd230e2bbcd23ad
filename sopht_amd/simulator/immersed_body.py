"""Host-side immersed-body glue (float64 numpy), feeding the device IB kernels.

Mirrors ``sopht/simulator/immersed_body``: ``ImmersedBodyForcingGrid``
(``immersed_body_forcing_grid.py:9-54``), ``ImmersedBodyFlowInteraction``
(``immersed_body_flow_interaction.py:15-141``), ``RigidBodyFlowInteraction``
(``rigid_body/rigid_body_flow_interaction.py:9-54``), ``ThreeDimensionalRigidBodyForcingGrid``
and ``SphereForcingGrid`` (``rigid_body/rigid_body_forcing_grids.py:108-300``). Only the
``(3, n)`` Lagrangian position and velocity arrays go to the device each call, and the
forcing comes back only when asked (``compute_flow_forces_and_torques``).

pyelastica is not a dependency: a rigid body is any object with ``position_collection``
(3, 1), ``director_collection`` (3, 3, 1), ``velocity_collection`` (3, 1) and
``omega_collection`` (3, 1), such as :class:`RigidBody` below.

Two forcing grids have no counterpart in the reference and serve the benchmark
workloads: :class:`FibonacciSphereForcingGrid` and :class:`FibonacciEllipsoidForcingGrid`
(Fibonacci-lattice surface points).
"""

from __future__ import annotations

import logging
from abc import abstractmethod

import numpy as np

from ..numeric.immersed_boundary_ops import VirtualBoundaryForcing
from ..utils import synthetic

logger = logging.getLogger(__name__)


def _batch_cross(a, b):
    return np.cross(a, b, axis=0)


class RigidBody:
    """Minimal rigid-body state with pyelastica's attribute layout."""

    def __init__(self, center, radius=0.0, velocity=(0.0, 0.0, 0.0), omega=(0.0, 0.0, 0.0)):
        self.position_collection = np.asarray(center, dtype=np.float64).reshape(3, 1).copy()
        self.director_collection = np.eye(3).reshape(3, 3, 1).copy()
        self.velocity_collection = np.asarray(velocity, dtype=np.float64).reshape(3, 1).copy()
        self.omega_collection = np.asarray(omega, dtype=np.float64).reshape(3, 1).copy()
        self.radius = float(radius)


class ImmersedBodyForcingGrid:
    """Base class of forcing grids (``immersed_body_forcing_grid.py:9-54``)."""

    def __init__(self, grid_dim: int, num_lag_nodes: int) -> None:
        self.grid_dim = grid_dim
        self.num_lag_nodes = num_lag_nodes
        self.position_field = np.zeros((self.grid_dim, self.num_lag_nodes))
        self.velocity_field = np.zeros_like(self.position_field)
        if grid_dim == 2:
            # the reference's notice for planar grids (immersed_body_forcing_grid.py:26-33)
            logger.warning(
                "\n=================================================="
                "\n2D body forcing grid generated, this assumes the body"
                "\nmoves in XY plane! Please initialize your body such that"
                "\nensuing dynamics are constrained in XY plane!"
                "\n=================================================="
            )

    @abstractmethod
    def compute_lag_grid_position_field(self) -> None: ...

    @abstractmethod
    def compute_lag_grid_velocity_field(self) -> None: ...

    @abstractmethod
    def transfer_forcing_from_grid_to_body(self, body_flow_forces, body_flow_torques, lag_grid_forcing_field) -> None: ...

    @abstractmethod
    def get_maximum_lagrangian_grid_spacing(self) -> float: ...


class ThreeDimensionalRigidBodyForcingGrid(ImmersedBodyForcingGrid):
    """``rigid_body_forcing_grids.py:108-173``."""

    def __init__(self, grid_dim, num_lag_nodes, rigid_body):
        if grid_dim != 3:
            msg = "Invalid grid dimensions. 3D Rigid body forcing grid is only defined for grid_dim=3"
            raise ValueError(msg)
        self.rigid_body = rigid_body
        super().__init__(grid_dim, num_lag_nodes)
        self.local_frame_relative_position_field = np.zeros_like(self.position_field)
        self.global_frame_relative_position_field = np.zeros_like(self.position_field)

    def compute_lag_grid_position_field(self):
        self.global_frame_relative_position_field[...] = np.dot(
            self.rigid_body.director_collection[:, :, 0].T, self.local_frame_relative_position_field
        )
        self.position_field[...] = self.rigid_body.position_collection + self.global_frame_relative_position_field

    def compute_lag_grid_velocity_field(self):
        # V + (omega x r) with omega broadcast over the nodes, written per component with the
        # operations np.cross performs (a1 b2 - a2 b1, a2 b0 - a0 b2, a0 b1 - a1 b0), so the
        # result is bit-identical to the reference's np.cross(omega * ones, r) at a fraction of
        # its cost (it sits between the CFL synchronisation and the next kernel launch).
        om = np.dot(self.rigid_body.director_collection[:, :, 0].T, self.rigid_body.omega_collection)[:, 0]
        r = self.global_frame_relative_position_field
        v = self.velocity_field
        v0 = self.rigid_body.velocity_collection[:, 0]
        v[0] = v0[0] + (om[1] * r[2] - om[2] * r[1])
        v[1] = v0[1] + (om[2] * r[0] - om[0] * r[2])
        v[2] = v0[2] + (om[0] * r[1] - om[1] * r[0])

    def transfer_forcing_from_grid_to_body(self, body_flow_forces, body_flow_torques, lag_grid_forcing_field):
        f = lag_grid_forcing_field
        if hasattr(f, "detach"):
            f = f.detach().double().cpu().numpy()
        body_flow_forces[...] = -np.sum(f, axis=1).reshape(-1, 1)
        body_flow_torques[...] = -np.dot(
            self.rigid_body.director_collection[:, :, 0],
            np.sum(_batch_cross(self.global_frame_relative_position_field, f), axis=1).reshape(-1, 1),
        )

    def get_maximum_lagrangian_grid_spacing(self) -> float:
        raise NotImplementedError


class SphereForcingGrid(ThreeDimensionalRigidBodyForcingGrid):
    """Latitude-ring sphere grid (``rigid_body_forcing_grids.py:236-300``)."""

    def __init__(self, grid_dim, rigid_body, num_forcing_points_along_equator):
        self.num_forcing_points_along_equator = num_forcing_points_along_equator
        polar = np.linspace(0, np.pi, num_forcing_points_along_equator // 2)
        per_lat = np.rint(num_forcing_points_along_equator * np.sin(polar)).astype(int) + 1
        super().__init__(grid_dim=grid_dim, num_lag_nodes=int(sum(per_lat)), rigid_body=rigid_body)
        xs, ys, zs = [], [], []
        r = rigid_body.radius
        for m, th in zip(per_lat, polar, strict=False):
            az = np.linspace(0.0, 2 * np.pi, m, endpoint=False)
            xs.append(r * np.sin(th) * np.cos(az))
            ys.append(r * np.sin(th) * np.sin(az))
            zs.append(r * np.cos(th) * np.ones(m))
        self.global_frame_relative_position_field[0] = np.concatenate(xs)
        self.global_frame_relative_position_field[1] = np.concatenate(ys)
        self.global_frame_relative_position_field[2] = np.concatenate(zs)
        self.compute_lag_grid_position_field()
        self.compute_lag_grid_velocity_field()

    def get_maximum_lagrangian_grid_spacing(self) -> float:
        return self.rigid_body.radius * (2 * np.pi / self.num_forcing_points_along_equator)

    def compute_lag_grid_position_field(self):
        self.position_field[...] = self.rigid_body.position_collection + self.global_frame_relative_position_field


class FibonacciEllipsoidForcingGrid(ThreeDimensionalRigidBodyForcingGrid):
    """Fibonacci-lattice points on an ellipsoid with semi-axes ``(a, b, c)`` in the body frame.

    The maximum Lagrangian spacing is the mean spacing ``sqrt(area / n)`` with the
    Knud Thomsen surface-area approximation (exact for a sphere)."""

    def __init__(self, grid_dim, rigid_body, num_lag_nodes, semi_axes):
        super().__init__(grid_dim=grid_dim, num_lag_nodes=num_lag_nodes, rigid_body=rigid_body)
        self.semi_axes = tuple(float(s) for s in semi_axes)
        self.local_frame_relative_position_field[...] = synthetic.fibonacci_ellipsoid(
            num_lag_nodes, self.semi_axes, (0.0, 0.0, 0.0)
        )
        self.compute_lag_grid_position_field()
        self.compute_lag_grid_velocity_field()

    def get_maximum_lagrangian_grid_spacing(self) -> float:
        a, b, c = self.semi_axes
        p = 1.6075
        area = 4 * np.pi * (((a * b) ** p + (a * c) ** p + (b * c) ** p) / 3) ** (1 / p)
        return float(np.sqrt(area / self.num_lag_nodes))


class FibonacciSphereForcingGrid(FibonacciEllipsoidForcingGrid):
    """Fibonacci-lattice sphere of the body's radius."""

    def __init__(self, grid_dim, rigid_body, num_lag_nodes):
        r = rigid_body.radius
        super().__init__(grid_dim, rigid_body, num_lag_nodes, (r, r, r))


class ImmersedBodyFlowInteraction(VirtualBoundaryForcing):
    """``immersed_body_flow_interaction.py:15-141``: VB forcing bound to the simulator's
    forcing and velocity fields (device tensors, held by reference)."""

    def __init__(
        self,
        eul_grid_forcing_field,
        eul_grid_velocity_field,
        body_flow_forces,
        body_flow_torques,
        forcing_grid_cls,
        virtual_boundary_stiffness_coeff,
        virtual_boundary_damping_coeff,
        dx,
        grid_dim,
        real_t=np.float64,
        eul_grid_coord_shift=None,
        interp_kernel_width=None,
        enable_eul_grid_forcing_reset=False,
        num_threads=False,
        start_time=0.0,
        **forcing_grid_kwargs,
    ):
        self.body_flow_forces = body_flow_forces
        self.body_flow_torques = body_flow_torques
        self.forcing_grid = forcing_grid_cls(grid_dim=grid_dim, **forcing_grid_kwargs)
        self.eul_grid_forcing_field = eul_grid_forcing_field
        self.eul_grid_velocity_field = eul_grid_velocity_field
        max_lag_grid_dx = self.forcing_grid.get_maximum_lagrangian_grid_spacing()
        self._log_resolution(max_lag_grid_dx, dx)
        # :83-86 rescale by the Lagrangian spacing
        virtual_boundary_stiffness_coeff *= max_lag_grid_dx ** (grid_dim - 1)
        virtual_boundary_damping_coeff *= max_lag_grid_dx ** (grid_dim - 1)
        super().__init__(
            virtual_boundary_stiffness_coeff,
            virtual_boundary_damping_coeff,
            grid_dim,
            dx,
            self.forcing_grid.num_lag_nodes,
            real_t,
            eul_grid_coord_shift,
            interp_kernel_width,
            enable_eul_grid_forcing_reset,
            num_threads,
            start_time,
        )

    def _log_resolution(self, max_lag_grid_dx, dx):
        """The reference's resolution notice, same text and levels (:52-79): a warning when
        the Lagrangian spacing is above 2 dx (the delta kernel's support) or below dx / 2."""
        banner = "\n=================================================="
        head = f"{banner}\nFor {type(self.forcing_grid).__name__}:\nEulerian grid spacing (dx): {dx}"
        if max_lag_grid_dx > 2 * dx:
            body = (f"\nMax Lagrangian grid spacing: {max_lag_grid_dx} > 2 * dx"
                    "\nThe Lagrangian grid of the body is too coarse relative to"
                    "\nthe Eulerian grid of the flow, which can lead to unexpected"
                    "\nconvergence. Please make the Lagrangian grid finer.")
            log = logger.warning
        elif max_lag_grid_dx < 0.5 * dx:
            body = (f"\nMax Lagrangian grid spacing: {max_lag_grid_dx} < 0.5 * dx"
                    "\nThe Lagrangian grid of the body is too fine relative to"
                    "\nthe Eulerian grid of the flow, which corresponds to redundant"
                    "\nforcing points. Please make the Lagrangian grid coarser.")
            log = logger.warning
        else:
            body = "\nLagrangian grid is resolved almost the same\nas the Eulerian grid of the flow."
            log = logger.info
        log(head + body + banner)

    def compute_interaction_on_lag_grid(self):
        self.forcing_grid.compute_lag_grid_position_field()
        self.forcing_grid.compute_lag_grid_velocity_field()
        self.compute_interaction_force_on_lag_grid(
            eul_grid_velocity_field=self.eul_grid_velocity_field,
            lag_grid_position_field=self.forcing_grid.position_field,
            lag_grid_velocity_field=self.forcing_grid.velocity_field,
        )

    def __call__(self):
        self.forcing_grid.compute_lag_grid_position_field()
        self.forcing_grid.compute_lag_grid_velocity_field()
        self.compute_interaction_forcing(
            eul_grid_forcing_field=self.eul_grid_forcing_field,
            eul_grid_velocity_field=self.eul_grid_velocity_field,
            lag_grid_position_field=self.forcing_grid.position_field,
            lag_grid_velocity_field=self.forcing_grid.velocity_field,
        )

    def compute_flow_forces_and_torques(self):
        self.compute_interaction_on_lag_grid()
        self.forcing_grid.transfer_forcing_from_grid_to_body(
            body_flow_forces=self.body_flow_forces,
            body_flow_torques=self.body_flow_torques,
            lag_grid_forcing_field=self.lag_grid_forcing_field,
        )

    def get_grid_deviation_error_l2_norm(self):
        import torch

        return float(torch.linalg.vector_norm(self.lag_grid_position_mismatch_field.double()).item()) / np.sqrt(
            self.forcing_grid.num_lag_nodes
        )


class RigidBodyFlowInteraction(ImmersedBodyFlowInteraction):
    """``rigid_body_flow_interaction.py:9-54``."""

    def __init__(
        self,
        rigid_body,
        eul_grid_forcing_field,
        eul_grid_velocity_field,
        virtual_boundary_stiffness_coeff,
        virtual_boundary_damping_coeff,
        dx,
        grid_dim,
        forcing_grid_cls,
        real_t=np.float64,
        eul_grid_coord_shift=None,
        interp_kernel_width=None,
        enable_eul_grid_forcing_reset=False,
        num_threads=False,
        start_time=0.0,
        **forcing_grid_kwargs,
    ):
        forcing_grid_kwargs["rigid_body"] = rigid_body
        super().__init__(
            eul_grid_forcing_field,
            eul_grid_velocity_field,
            np.zeros((3, 1)),
            np.zeros((3, 1)),
            forcing_grid_cls,
            virtual_boundary_stiffness_coeff,
            virtual_boundary_damping_coeff,
            dx,
            grid_dim,
            real_t,
            eul_grid_coord_shift,
            interp_kernel_width,
            enable_eul_grid_forcing_reset,
            num_threads,
            start_time,
            **forcing_grid_kwargs,
        )
