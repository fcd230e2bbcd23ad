"""Host-side rigid-body forcing grid: node velocities V + omega x r must equal the reference's
np.cross(omega * ones, r) bit for bit (rigid_body_forcing_grids.py:108-173)."""

import numpy as np

from sopht_amd.simulator.immersed_body import FibonacciSphereForcingGrid, RigidBody


def test_rigid_grid_velocity_matches_np_cross_bitwise():
    rng = np.random.default_rng(7)
    body = RigidBody((0.5, 0.4, 0.6), 0.08, velocity=rng.standard_normal(3), omega=rng.standard_normal(3))
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    body.director_collection[:, :, 0] = q
    grid = FibonacciSphereForcingGrid(3, body, 1000)
    grid.compute_lag_grid_position_field()
    grid.compute_lag_grid_velocity_field()
    omega = np.dot(body.director_collection[:, :, 0].T, body.omega_collection)
    expected = body.velocity_collection + np.cross(
        omega * np.ones(grid.num_lag_nodes), grid.global_frame_relative_position_field, axis=0
    )
    assert np.array_equal(expected, grid.velocity_field)


def test_rigid_grid_position_is_body_frame_rotation_plus_centre():
    body = RigidBody((0.25, 0.5, 0.75), 0.1)
    rot = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    body.director_collection[:, :, 0] = rot
    grid = FibonacciSphereForcingGrid(3, body, 200)
    grid.compute_lag_grid_position_field()
    rel = rot.T @ grid.local_frame_relative_position_field
    assert np.allclose(grid.position_field, body.position_collection + rel, rtol=0, atol=1e-15)
    # a sphere's nodes sit on its surface
    assert np.allclose(np.linalg.norm(grid.position_field - body.position_collection, axis=0), 0.1)
