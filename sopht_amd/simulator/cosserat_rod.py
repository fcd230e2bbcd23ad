"""Cosserat-rod coupling glue: forcing grids on a rod, the rod-flow interactor and the
pyelastica forcing hook.

Mirrors ``sopht/simulator/immersed_body/cosserat_rod/cosserat_rod_forcing_grids.py:10-589``
(``CosseratRodNodalForcingGrid``, ``CosseratRodElementCentricForcingGrid``,
``CosseratRodSurfaceForcingGrid``), ``cosserat_rod_flow_interaction.py:9-58``
(``CosseratRodFlowInteraction``) and ``immersed_body/flow_forces.py:12-25`` (``FlowForces``).
The rod itself stays on the host (pyelastica, or any object with its attribute layout:
``n_elems``, ``position_collection`` (3, n+1), ``velocity_collection`` (3, n+1),
``director_collection`` (3, 3, n), ``omega_collection`` (3, n), ``lengths`` (n),
``radius`` (n), ``mass`` (n+1)); only the grid's (3, n_lag) positions and velocities go to
the device, and the Lagrangian forcing comes back for the force transfer.

pyelastica is not a dependency. The small array helpers it provides to the reference
(batched cross products and mat-vecs, node/element averaging) are written here from their
definitions, and :class:`StraightRod` builds a straight rod with pyelastica's attribute
layout (``CosseratRod.straight_rod``'s geometry) for rods whose motion is prescribed on
the host. The 2D-only ``CosseratRodEdgeForcingGrid`` is outside this 3D backend.
"""

from __future__ import annotations

import numpy as np

from .immersed_body import ImmersedBodyFlowInteraction, ImmersedBodyForcingGrid


# ----------------------------------------------------------------------------- array helpers


def batch_cross(a, b):
    """Cross product of matching columns of two (3, n) arrays."""
    out = np.empty(np.broadcast_shapes(np.shape(a), np.shape(b)))
    out[0] = a[1] * b[2] - a[2] * b[1]
    out[1] = a[2] * b[0] - a[0] * b[2]
    out[2] = a[0] * b[1] - a[1] * b[0]
    return out


def batch_matvec(mats, vecs):
    """``out[:, k] = mats[:, :, k] @ vecs[:, k]`` for (3, 3, n) matrices and (3, n) vectors."""
    return np.einsum("ijk,jk->ik", mats, vecs)


def batch_matrix_transpose(mats):
    """Transpose of each (3, 3) slice of a (3, 3, n) stack."""
    return np.ascontiguousarray(np.transpose(mats, (1, 0, 2)))


def node_to_element_velocity(mass, node_velocity):
    """Momentum-conserving element velocity: the mass-weighted mean of the element's two
    nodes, ``(m_i v_i + m_{i+1} v_{i+1}) / (m_i + m_{i+1})``."""
    m = np.asarray(mass, dtype=np.float64)
    return (node_velocity[:, :-1] * m[:-1] + node_velocity[:, 1:] * m[1:]) / (m[:-1] + m[1:])


def elements_to_nodes_inplace(element_vector, node_vector):
    """Half of each element's vector to each of its two nodes (added in place)."""
    node_vector[:, :-1] += 0.5 * element_vector
    node_vector[:, 1:] += 0.5 * element_vector


def _host(field):
    """A (grid_dim, n) Lagrangian field as float64 host memory (device tensors are copied)."""
    if hasattr(field, "detach"):
        return field.detach().double().cpu().numpy()
    return np.asarray(field, dtype=np.float64)


# ----------------------------------------------------------------------------- a straight rod


class StraightRod:
    """A straight rod with pyelastica's ``CosseratRod`` attribute layout.

    Geometry as ``CosseratRod.straight_rod(n_elements, start, direction, normal,
    base_length, base_radius, density)``: nodes evenly spaced from ``start`` along the unit
    ``direction``; every element's directors are the rows ``d1 = normal``,
    ``d2 = direction x normal``, ``d3 = direction``; element mass ``density * pi r^2 * l``
    split half to each node. ``base_radius`` may be a scalar or one value per element.
    The state (positions, velocities, directors, angular velocities) is plain arrays the
    caller moves; nothing here integrates the rod's mechanics."""

    def __init__(self, n_elements, start, direction, normal, base_length, base_radius, density=1.0):
        n = int(n_elements)
        d3 = np.asarray(direction, dtype=np.float64)
        d3 = d3 / np.linalg.norm(d3)
        d1 = np.asarray(normal, dtype=np.float64)
        d1 = d1 / np.linalg.norm(d1)
        if abs(float(d1 @ d3)) > 1e-12:
            msg = "normal must be perpendicular to the direction"
            raise ValueError(msg)
        d2 = np.cross(d3, d1)
        self.n_elems = n
        s = np.linspace(0.0, float(base_length), n + 1)
        self.position_collection = np.asarray(start, dtype=np.float64).reshape(3, 1) + d3.reshape(3, 1) * s
        self.velocity_collection = np.zeros((3, n + 1))
        self.director_collection = np.repeat(np.stack([d1, d2, d3])[:, :, None], n, axis=2)
        self.omega_collection = np.zeros((3, n))
        self.radius = np.broadcast_to(np.asarray(base_radius, dtype=np.float64), (n,)).copy()
        self.lengths = np.full(n, float(base_length) / n)
        self.tangents = np.repeat(d3.reshape(3, 1), n, axis=1)
        elem_mass = float(density) * np.pi * self.radius**2 * self.lengths
        self.mass = np.zeros(n + 1)
        self.mass[:-1] += 0.5 * elem_mass
        self.mass[1:] += 0.5 * elem_mass
        self.external_forces = np.zeros((3, n + 1))
        self.external_torques = np.zeros((3, n))


# ----------------------------------------------------------------------------- forcing grids


class CosseratRodNodalForcingGrid(ImmersedBodyForcingGrid):
    """Forcing points at the rod's nodes (``cosserat_rod_forcing_grids.py:10-79``)."""

    def __init__(self, grid_dim, cosserat_rod):
        super().__init__(grid_dim, cosserat_rod.n_elems + 1)
        self.cosserat_rod = cosserat_rod
        self.moment_arm = np.zeros((3, cosserat_rod.n_elems))
        self.compute_lag_grid_position_field()
        self.compute_lag_grid_velocity_field()

    def compute_lag_grid_position_field(self):
        self.position_field[...] = self.cosserat_rod.position_collection[: self.grid_dim]

    def compute_lag_grid_velocity_field(self):
        self.velocity_field[...] = self.cosserat_rod.velocity_collection[: self.grid_dim]

    def transfer_forcing_from_grid_to_body(self, body_flow_forces, body_flow_torques, lag_grid_forcing_field):
        rod = self.cosserat_rod
        # each node carries minus its forcing point's force (action and reaction)
        body_flow_forces[: self.grid_dim] = -_host(lag_grid_forcing_field)
        f = body_flow_forces
        # element torque: the half-element arms times the force differences, plus the end
        # nodes' own halves
        half_edge = 0.5 * (rod.position_collection[:, 1:] - rod.position_collection[:, :-1])
        self.moment_arm[...] = half_edge
        tau = batch_cross(half_edge, 0.5 * (f[:, 1:] - f[:, :-1]))
        tau[:, -1] += 0.5 * np.cross(half_edge[:, -1], f[:, -1])
        tau[:, 0] -= 0.5 * np.cross(half_edge[:, 0], f[:, 0])
        # into each element's material frame
        body_flow_torques[...] = batch_matvec(rod.director_collection, tau)

    def get_maximum_lagrangian_grid_spacing(self):
        return np.amax(self.cosserat_rod.lengths)


class CosseratRodElementCentricForcingGrid(ImmersedBodyForcingGrid):
    """Forcing points at element centres (``cosserat_rod_forcing_grids.py:82-131``)."""

    def __init__(self, grid_dim, cosserat_rod):
        super().__init__(grid_dim, cosserat_rod.n_elems)
        self.cosserat_rod = cosserat_rod
        self.compute_lag_grid_position_field()
        self.compute_lag_grid_velocity_field()

    def compute_lag_grid_position_field(self):
        x = self.cosserat_rod.position_collection
        self.position_field[...] = 0.5 * (x[: self.grid_dim, 1:] + x[: self.grid_dim, :-1])

    def compute_lag_grid_velocity_field(self):
        rod = self.cosserat_rod
        self.velocity_field[...] = node_to_element_velocity(rod.mass, rod.velocity_collection)[: self.grid_dim]

    def transfer_forcing_from_grid_to_body(self, body_flow_forces, body_flow_torques, lag_grid_forcing_field):
        # half of each element point's reaction to each of its nodes; no torque at centres
        body_flow_forces[...] = 0.0
        reaction = np.zeros((body_flow_forces.shape[0], self.num_lag_nodes))
        reaction[: self.grid_dim] = -_host(lag_grid_forcing_field)
        elements_to_nodes_inplace(reaction, body_flow_forces)

    def get_maximum_lagrangian_grid_spacing(self):
        return np.amax(self.cosserat_rod.lengths)


class CosseratRodSurfaceForcingGrid(ImmersedBodyForcingGrid):
    """Forcing points on rings around each element's centre, in its d1-d2 plane
    (``cosserat_rod_forcing_grids.py:292-589``).

    Element i gets ``rint(r_i / max r * density)`` points on its circumference (one point at
    its centre when that is below 3). With ``with_cap``, each end element also gets
    concentric inner rings that close the rod's end: ``k = max(int(r // (r 2 pi / m)), 1)``
    rings of ``int(linspace(1, m, k, endpoint=False))`` points at radius fractions
    ``j / k``, stored after the element's outer ring."""

    def __init__(self, grid_dim, cosserat_rod, surface_grid_density_for_largest_element, with_cap=False):
        if grid_dim != 3:
            msg = "Invalid grid dimensions. Cosserat rod surface forcing grid is only defined for grid_dim=3"
            raise ValueError(msg)
        rod = cosserat_rod
        self.cosserat_rod = rod
        self.n_elems = rod.n_elems
        self.surface_grid_density_for_largest_element = surface_grid_density_for_largest_element
        self.with_cap = with_cap
        radius = np.asarray(rod.radius, dtype=np.float64)
        outer = np.rint(radius / np.max(radius) * surface_grid_density_for_largest_element).astype(int)
        outer[outer < 3] = 1
        # per element: the angle of each point and its radius fraction
        angles, fractions = [], []
        for i in range(self.n_elems):
            m = int(outer[i])
            if m > 1:
                a = np.linspace(0.0, 2 * np.pi, m, endpoint=False)
            else:
                a = np.array([])  # a single point at the centre
            angles.append(a)
            fractions.append(np.ones(m))
        if with_cap:
            for e in (0, self.n_elems - 1):
                m = int(outer[e])
                if m <= 1:
                    continue
                r = radius[e]
                rings = max(int(r // (r * (2.0 * np.pi / m))), 1)
                counts = np.linspace(1, m, rings, endpoint=False).astype(int)
                angles[e] = np.concatenate(
                    [angles[e]] + [np.linspace(0.0, 2 * np.pi, c, endpoint=False) for c in counts]
                )
                ring_frac = np.linspace(0.0, r, rings, endpoint=False) / r
                fractions[e] = np.concatenate([fractions[e]] + [np.full(c, f) for c, f in zip(counts, ring_frac)])
        self.surface_grid_points = np.array([f.size for f in fractions], dtype=int)
        self.surface_point_rotation_angle_list = angles
        self.grid_point_radius_ratio = np.concatenate(fractions)
        super().__init__(grid_dim, int(self.surface_grid_points.sum()))
        self.end_idx = np.cumsum(self.surface_grid_points).astype(int)
        self.start_idx = (self.end_idx - self.surface_grid_points).astype(int)
        self.element_of_point = np.repeat(np.arange(self.n_elems), self.surface_grid_points)
        self.local_frame_surface_points = np.zeros_like(self.position_field)
        for i, a in enumerate(angles):
            if a.size:
                sl = slice(self.start_idx[i], self.end_idx[i])
                self.local_frame_surface_points[0, sl] = np.cos(a)
                self.local_frame_surface_points[1, sl] = np.sin(a)
        self.moment_arm = np.zeros_like(self.position_field)
        self.grid_point_radius = np.zeros(self.num_lag_nodes)
        self.lag_grid_torque_field = np.zeros_like(self.position_field)
        self.compute_lag_grid_position_field()
        self.compute_lag_grid_velocity_field()

    def compute_lag_grid_position_field(self):
        rod = self.cosserat_rod
        e = self.element_of_point
        x = rod.position_collection
        centre = 0.5 * (x[:, 1:] + x[:, :-1])
        self.grid_point_radius[...] = np.asarray(rod.radius, dtype=np.float64)[e] * self.grid_point_radius_ratio
        # Q^T (lab frame) of each point's element applied to its in-plane offset
        qt = batch_matrix_transpose(rod.director_collection)[:, :, e]
        self.moment_arm[...] = self.grid_point_radius * batch_matvec(qt, self.local_frame_surface_points)
        self.position_field[...] = centre[:, e] + self.moment_arm

    def compute_lag_grid_velocity_field(self):
        rod = self.cosserat_rod
        e = self.element_of_point
        v_elem = node_to_element_velocity(rod.mass, rod.velocity_collection)
        omega_lab = batch_matvec(batch_matrix_transpose(rod.director_collection), rod.omega_collection)
        self.velocity_field[...] = v_elem[:, e] + batch_cross(omega_lab[:, e], self.moment_arm)

    def transfer_forcing_from_grid_to_body(self, body_flow_forces, body_flow_torques, lag_grid_forcing_field):
        rod = self.cosserat_rod
        f = _host(lag_grid_forcing_field)
        # per element sums of the point forces and of their torques about the element centre
        per_elem = np.add.reduceat(f, self.start_idx, axis=1) if self.num_lag_nodes else np.zeros((3, self.n_elems))
        body_flow_forces[...] = 0.0
        body_flow_forces[:, :-1] -= 0.5 * per_elem
        body_flow_forces[:, 1:] -= 0.5 * per_elem
        self.lag_grid_torque_field[...] = batch_cross(self.moment_arm, -f)
        tau = np.add.reduceat(self.lag_grid_torque_field, self.start_idx, axis=1)
        body_flow_torques[...] = batch_matvec(rod.director_collection, tau)

    def get_maximum_lagrangian_grid_spacing(self):
        rod = self.cosserat_rod
        arc = np.asarray(rod.radius, dtype=np.float64) * (2 * np.pi / self.surface_grid_density_for_largest_element)
        return float(np.amax([np.asarray(rod.lengths, dtype=np.float64), arc]))


# ----------------------------------------------------------------------------- interaction


class CosseratRodFlowInteraction(ImmersedBodyFlowInteraction):
    """``cosserat_rod_flow_interaction.py:9-58``: the interactor of a rod, with nodal body
    forces (3, n+1) and element torques (3, n)."""

    def __init__(
        self,
        cosserat_rod,
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
        forcing_grid_kwargs["cosserat_rod"] = cosserat_rod
        super().__init__(
            eul_grid_forcing_field,
            eul_grid_velocity_field,
            np.zeros((3, cosserat_rod.n_elems + 1)),
            np.zeros((3, cosserat_rod.n_elems)),
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


try:  # a pyelastica forcing must derive from NoForces; without pyelastica, a plain base
    from elastica import NoForces as _ForcingBase
except ImportError:  # pragma: no cover - pyelastica is optional
    _ForcingBase = object


class FlowForces(_ForcingBase):
    """The flow's forcing on a body, as a pyelastica forcing (``flow_forces.py:12-25``):
    each call evaluates the interactor's forces and torques on the body (interpolation and
    Lagrangian forcing on the device, no spreading) and adds them to the system's external
    loads. With pyelastica installed it derives from ``NoForces``, so
    ``add_forcing_to(rod).using(FlowForces, interactor)`` accepts it."""

    def __init__(self, body_flow_interactor):
        if _ForcingBase is not object:
            super().__init__()
        self.body_flow_interactor = body_flow_interactor

    def apply_forces(self, system, time=0.0):
        self.body_flow_interactor.compute_flow_forces_and_torques()
        system.external_forces += self.body_flow_interactor.body_flow_forces
        system.external_torques += self.body_flow_interactor.body_flow_torques

    def apply_torques(self, system, time=0.0):
        """Torques are added with the forces in :meth:`apply_forces`."""
