"""Host-side Cosserat-rod glue, on the reference's own cases
(``tests/test_simulator/test_immersed_body/cosserat_rod/test_cosserat_rod_forcing_grids.py:9-753``,
``test_flow_forces.py:5-30``): nodal, element-centric and surface forcing grids on a straight
rod along (1, 1, 1) with linearly increasing node velocities and element angular velocities,
their force and torque transfer under uniform forcing, and the pyelastica forcing hook with
mock objects. pyelastica is absent, so the rod is :class:`StraightRod` with the geometry of
``CosseratRod.straight_rod``; the expected values below are built independently from the
closed forms the reference's tests state."""

import numpy as np
import pytest

import sopht_amd.simulator as sps
from sopht_amd.simulator import cosserat_rod as cr

TOL = 1e3 * np.finfo(np.float64).eps


def mock_straight_rod(n_elems, base_radius=0.05):
    """The reference tests' rod: start 0, direction (1, 1, 1), normal (0, -1, 1), length 1,
    density 1e3, node velocities 1..n+1 in every component, element omegas 1..n."""
    rod = sps.StraightRod(n_elems, np.zeros(3), np.array([1.0, 1.0, 1.0]), np.array([0.0, -1.0, 1.0]), 1.0,
                          base_radius, density=1e3)
    rod.velocity_collection[...] = np.linspace(1, n_elems + 1, n_elems + 1)
    rod.omega_collection[...] = np.linspace(1, n_elems, n_elems)
    return rod


@pytest.mark.parametrize("n_elems", [8, 16])
def test_straight_rod_geometry(n_elems):
    rod = mock_straight_rod(n_elems)
    t = np.ones(3) / np.sqrt(3)
    np.testing.assert_allclose(rod.position_collection[:, -1], t, atol=TOL)
    for k in range(n_elems):
        q = rod.director_collection[:, :, k]
        np.testing.assert_allclose(q @ q.T, np.eye(3), atol=TOL)
        np.testing.assert_allclose(np.linalg.det(q), 1.0, atol=TOL)  # right-handed
        np.testing.assert_allclose(q[2], t, atol=TOL)  # d3 is the tangent
    np.testing.assert_allclose(rod.lengths, 1.0 / n_elems)
    # lumped masses: half elements at the ends, whole ones inside
    m_e = 1e3 * np.pi * 0.05**2 / n_elems
    np.testing.assert_allclose(rod.mass[[0, -1]], 0.5 * m_e)
    np.testing.assert_allclose(rod.mass[1:-1], m_e)


@pytest.mark.parametrize("n_elems", [8, 16])
def test_node_to_element_velocity(n_elems):
    """``test_cosserat_rod_forcing_grids.py:33-55``: averages inside, (v0 + 2 v1)/3 at the ends
    (half-mass end nodes)."""
    rod = mock_straight_rod(n_elems)
    v = rod.velocity_collection
    got = cr.node_to_element_velocity(rod.mass, v)
    want = 0.5 * (v[:, 1:] + v[:, :-1])
    want[:, 0] = (v[:, 0] + 2 * v[:, 1]) / 3
    want[:, -1] = (v[:, -1] + 2 * v[:, -2]) / 3
    np.testing.assert_allclose(got, want)


@pytest.mark.parametrize("n_elems", [8, 16])
def test_elements_to_nodes_inplace(n_elems, rng):
    """``test_cosserat_rod_forcing_grids.py:57-72``."""
    e = rng.random((3, n_elems))
    nodes = np.zeros((3, n_elems + 1))
    cr.elements_to_nodes_inplace(e, nodes)
    want = np.zeros((3, n_elems + 1))
    want[:, 1:] += 0.5 * e
    want[:, :-1] += 0.5 * e
    np.testing.assert_allclose(nodes, want)


def test_batch_helpers(rng):
    a, b = rng.random((3, 7)), rng.random((3, 7))
    np.testing.assert_allclose(cr.batch_cross(a, b), np.cross(a.T, b.T).T, atol=TOL)
    m = rng.random((3, 3, 7))
    np.testing.assert_allclose(cr.batch_matvec(m, a), np.stack([m[:, :, k] @ a[:, k] for k in range(7)], 1))
    np.testing.assert_array_equal(cr.batch_matrix_transpose(m)[:, :, 3], m[:, :, 3].T)


# ----------------------------------------------------------------------------- nodal grid


@pytest.mark.parametrize("grid_dim", [2, 3])
@pytest.mark.parametrize("n_elems", [8, 16])
def test_rod_nodal_grid_kinematics(grid_dim, n_elems):
    """``:74-97``."""
    rod = mock_straight_rod(n_elems)
    g = sps.CosseratRodNodalForcingGrid(grid_dim=grid_dim, cosserat_rod=rod)
    assert g.cosserat_rod is rod
    assert g.num_lag_nodes == n_elems + 1
    assert g.position_field.shape == (grid_dim, n_elems + 1)
    np.testing.assert_allclose(g.position_field, rod.position_collection[:grid_dim])
    np.testing.assert_allclose(g.velocity_field, rod.velocity_collection[:grid_dim])


@pytest.mark.parametrize("grid_dim", [2, 3])
@pytest.mark.parametrize("n_elems", [8, 16])
def test_rod_nodal_grid_force_transfer(grid_dim, n_elems):
    """``:99-141``: uniform forcing gives minus it at every node and, the inner torques
    cancelling, only the two end-element corrections."""
    rod = mock_straight_rod(n_elems)
    g = sps.CosseratRodNodalForcingGrid(grid_dim=grid_dim, cosserat_rod=rod)
    forces = np.zeros((3, n_elems + 1))
    torques = np.zeros((3, n_elems))
    f = np.zeros((grid_dim, g.num_lag_nodes))
    uniform = np.linspace(1.0, grid_dim, grid_dim).reshape(-1, 1)
    f[...] = uniform
    g.transfer_forcing_from_grid_to_body(body_flow_forces=forces, body_flow_torques=torques, lag_grid_forcing_field=f)
    want_f = np.zeros_like(forces)
    want_f[:grid_dim] = -uniform
    np.testing.assert_allclose(forces, want_f)
    arm = 0.5 * (rod.position_collection[:, 1:] - rod.position_collection[:, :-1])
    want_t = np.zeros_like(torques)
    want_t[:, -1] += rod.director_collection[:, :, -1] @ (np.cross(arm[:, -1], want_f[:, -1]) / 2.0)
    want_t[:, 0] -= rod.director_collection[:, :, 0] @ (np.cross(arm[:, 0], want_f[:, 0]) / 2.0)
    np.testing.assert_allclose(torques, want_t, atol=TOL)


@pytest.mark.parametrize("n_elems", [8, 16])
def test_rod_nodal_grid_spacing(n_elems):
    rod = mock_straight_rod(n_elems)
    g = sps.CosseratRodNodalForcingGrid(grid_dim=3, cosserat_rod=rod)
    assert g.get_maximum_lagrangian_grid_spacing() == rod.lengths[0]


# ----------------------------------------------------------------------------- element-centric grid


@pytest.mark.parametrize("grid_dim", [2, 3])
@pytest.mark.parametrize("n_elems", [8, 16])
def test_rod_element_centric_grid_kinematics(grid_dim, n_elems):
    """``:153-195``."""
    rod = mock_straight_rod(n_elems)
    g = sps.CosseratRodElementCentricForcingGrid(grid_dim=grid_dim, cosserat_rod=rod)
    assert g.num_lag_nodes == n_elems
    x, v = rod.position_collection, rod.velocity_collection
    start, end = x[:, :2].mean(axis=1), x[:, -2:].mean(axis=1)
    want = np.stack([np.linspace(start[a], end[a], n_elems) for a in range(grid_dim)])
    np.testing.assert_allclose(g.position_field, want)
    vs, ve = v[:, :2].mean(axis=1), v[:, -2:].mean(axis=1)
    want_v = np.stack([np.linspace(vs[a], ve[a], n_elems) for a in range(grid_dim)])
    want_v[:, 0] = (v[:grid_dim, 0] + 2 * v[:grid_dim, 1]) / 3
    want_v[:, -1] = (v[:grid_dim, -1] + 2 * v[:grid_dim, -2]) / 3
    np.testing.assert_allclose(g.velocity_field, want_v)


@pytest.mark.parametrize("grid_dim", [2, 3])
@pytest.mark.parametrize("n_elems", [8, 16])
def test_rod_element_centric_grid_force_transfer(grid_dim, n_elems):
    """``:197-224``: minus the forcing at inner nodes, half of it at the two ends, no torque."""
    rod = mock_straight_rod(n_elems)
    g = sps.CosseratRodElementCentricForcingGrid(grid_dim=grid_dim, cosserat_rod=rod)
    forces, torques = np.zeros((3, n_elems + 1)), np.zeros((3, n_elems))
    uniform = np.linspace(1.0, grid_dim, grid_dim).reshape(-1, 1)
    f = np.zeros((grid_dim, g.num_lag_nodes)) + uniform
    g.transfer_forcing_from_grid_to_body(body_flow_forces=forces, body_flow_torques=torques, lag_grid_forcing_field=f)
    want = np.zeros_like(forces)
    want[:grid_dim] = -uniform
    want[:grid_dim, (0, -1)] *= 0.5
    np.testing.assert_allclose(forces, want)
    np.testing.assert_allclose(torques, 0.0)


@pytest.mark.parametrize("n_elems", [8, 16])
def test_rod_element_centric_grid_spacing(n_elems):
    rod = mock_straight_rod(n_elems)
    g = sps.CosseratRodElementCentricForcingGrid(grid_dim=3, cosserat_rod=rod)
    assert g.get_maximum_lagrangian_grid_spacing() == rod.lengths[0]


# ----------------------------------------------------------------------------- surface grid


def expected_surface_layout(n_elems, density, radius, with_cap):
    """Per element: point count, angles and radius fractions, following the reference
    test's construction (``:451-547``): ``round(r_i / max r * density)`` outer points (one
    centre point below 3); with caps, ``k = max(int(r // (r 2 pi / m)), 1)`` inner rings of
    ``(m // k - 1) j + 1`` points at fractions ``j / k``."""
    counts, angles, fracs = [], [], []
    for i in range(n_elems):
        m = round(radius[i] / np.max(radius) * density)
        if m < 3:
            counts.append(1)
            angles.append(np.array([]))
            fracs.append(np.ones(1))
            continue
        a = [2 * np.pi / m * j for j in range(m)]
        fr = [1.0] * m
        total = m
        if with_cap and i in (0, n_elems - 1):
            k = max(int(radius[i] // (radius[i] * 2.0 * np.pi / m)), 1)
            rings = [(m // k - 1) * j + 1 for j in range(k)]
            for j, c in enumerate(rings):
                a += list(np.linspace(0, 2 * np.pi, c, endpoint=False))
                fr += [j / k] * c
            total += sum(rings)
        counts.append(total)
        angles.append(np.array(a))
        fracs.append(np.array(fr))
    return np.array(counts), angles, fracs


def test_rod_surface_grid_dimension():
    rod = mock_straight_rod(8)
    for d in (0, 1, 2, 4):
        with pytest.raises(ValueError, match="only defined for grid_dim=3"):
            sps.CosseratRodSurfaceForcingGrid(grid_dim=d, cosserat_rod=rod, surface_grid_density_for_largest_element=1)


SURFACE_CASES = [(n, d, t, c) for n in (8, 16) for d in (16, 12, 8, 4) for t in (1, 2, 5, 10) for c in (True, False)]


@pytest.mark.parametrize(("n_elems", "density", "taper", "with_cap"), SURFACE_CASES)
def test_rod_surface_grid_setup_kinematics_and_transfer(n_elems, density, taper, with_cap):
    """``:565-728``: layout, positions (centre + r Q^T local offset), velocities
    (element velocity + omega x arm), and uniform-force transfer (half of each element's
    total to each node, zero torque)."""
    radius = np.linspace(1, 1 / taper, n_elems)
    rod = mock_straight_rod(n_elems, base_radius=radius)
    g = sps.CosseratRodSurfaceForcingGrid(grid_dim=3, cosserat_rod=rod, surface_grid_density_for_largest_element=density,
                                          with_cap=with_cap)
    counts, angles, fracs = expected_surface_layout(n_elems, density, radius, with_cap)
    np.testing.assert_array_equal(g.surface_grid_points, counts)
    assert g.num_lag_nodes == counts.sum()
    end = np.cumsum(counts)
    np.testing.assert_array_equal(g.start_idx, end - counts)
    np.testing.assert_array_equal(g.end_idx, end)
    for i in range(n_elems):
        np.testing.assert_allclose(g.surface_point_rotation_angle_list[i], angles[i], atol=TOL)
    # positions and velocities point by point
    x, v, m = rod.position_collection, rod.velocity_collection, rod.mass
    p = 0
    for i in range(n_elems):
        centre = 0.5 * (x[:, i] + x[:, i + 1])
        qt = rod.director_collection[:, :, i].T
        v_el = (v[:, i] * m[i] + v[:, i + 1] * m[i + 1]) / (m[i] + m[i + 1])
        om = qt @ rod.omega_collection[:, i]
        for j in range(counts[i]):
            local = np.zeros(3) if angles[i].size == 0 else np.array([np.cos(angles[i][j]), np.sin(angles[i][j]), 0.0])
            arm = radius[i] * fracs[i][j] * (qt @ local)
            np.testing.assert_allclose(g.moment_arm[:, p], arm, atol=TOL)
            np.testing.assert_allclose(g.position_field[:, p], centre + arm, atol=TOL)
            np.testing.assert_allclose(g.velocity_field[:, p], v_el + np.cross(om, arm), atol=10 * TOL)
            p += 1
    forces, torques = np.zeros((3, n_elems + 1)), np.zeros((3, n_elems))
    uniform = np.linspace(1.0, 3, 3).reshape(-1, 1)
    g.transfer_forcing_from_grid_to_body(forces, torques, np.zeros((3, g.num_lag_nodes)) + uniform)
    want = np.zeros_like(forces)
    for i in range(n_elems):
        want[:, i] -= 0.5 * uniform[:, 0] * counts[i]
        want[:, i + 1] -= 0.5 * uniform[:, 0] * counts[i]
    np.testing.assert_allclose(forces, want)
    np.testing.assert_allclose(torques, 0.0, atol=TOL * max(1, density))


@pytest.mark.parametrize("n_elems", [8, 16])
@pytest.mark.parametrize("density", [16, 12, 8, 4])
@pytest.mark.parametrize("taper", [1, 2, 5, 10])
def test_rod_surface_grid_spacing(n_elems, density, taper):
    """``:730-753``."""
    radius = np.linspace(1, 1 / taper, n_elems)
    rod = mock_straight_rod(n_elems, base_radius=radius)
    g = sps.CosseratRodSurfaceForcingGrid(grid_dim=3, cosserat_rod=rod, surface_grid_density_for_largest_element=density)
    np.testing.assert_allclose(g.get_maximum_lagrangian_grid_spacing(),
                               max(rod.lengths[0], np.max(radius) * (2 * np.pi / density)))


def test_rod_surface_grid_torque_of_tangential_forcing():
    """A ring of forcing points each pushed along its own tangential direction d(theta) x d3
    (a pure twist): the element torque in the material frame is about d3 alone and equals
    the sum of r |f| over the ring (sign from the reaction)."""
    n = 2
    rod = sps.StraightRod(n, np.zeros(3), np.array([0, 0, 1.0]), np.array([1.0, 0, 0]), 1.0, 0.1)
    g = sps.CosseratRodSurfaceForcingGrid(grid_dim=3, cosserat_rod=rod, surface_grid_density_for_largest_element=8)
    d3 = np.array([0, 0, 1.0])
    f = np.cross(np.tile(d3, (g.num_lag_nodes, 1)), g.moment_arm.T).T  # d3 x arm: tangential, |f| = r
    forces, torques = np.zeros((3, n + 1)), np.zeros((3, n))
    g.transfer_forcing_from_grid_to_body(forces, torques, f)
    np.testing.assert_allclose(forces, 0.0, atol=TOL)
    np.testing.assert_allclose(torques[:2], 0.0, atol=TOL)
    # reaction -f: arm x (-(d3 x arm)) = -r^2 d3 per point, 8 points per element
    np.testing.assert_allclose(torques[2], -8 * 0.1**2, atol=TOL)


# ----------------------------------------------------------------------------- FlowForces


class MockBodyFlowInteractor:
    def __init__(self):
        self.body_flow_forces = 0.0
        self.body_flow_torques = 0.0

    def compute_flow_forces_and_torques(self):
        self.body_flow_forces = 1.0
        self.body_flow_torques = 2.0


class MockRod:
    def __init__(self):
        self.external_forces = 3.0
        self.external_torques = 4.0


def test_flow_forces():
    """``test_flow_forces.py:5-30``."""
    inter = MockBodyFlowInteractor()
    ff = sps.FlowForces(body_flow_interactor=inter)
    assert ff.body_flow_interactor is inter
    rod = MockRod()
    ff.apply_forces(system=rod)
    np.testing.assert_allclose(rod.external_forces, 4.0)
    np.testing.assert_allclose(rod.external_torques, 6.0)
