"""Pins the oracle's composite flow step and stable time step (CPU only).

The composite step is checked against the manual kernel sequence of the reference's
integration test (``tests/test_simulator/test_flow/test_navier_stokes_flow_simulators.py:163-308``)
built from the separately pinned pieces, and ``dt`` against the closed form of ``:311-338``.
"""

import numpy as np
import pytest

from oracle import flow, poisson, stencils


@pytest.mark.parametrize("n", [4, 8])
@pytest.mark.parametrize("real_t", [np.float32, np.float64])
@pytest.mark.parametrize("with_free_stream", [True, False])
@pytest.mark.parametrize("rho", [1.0, 0.1])
def test_step_equals_manual_sequence(n, real_t, with_free_stream, rho, rng):
    nu = 1e-2
    dt = 2.0
    U = np.array([3.0, 4.0, 5.0])
    sim = flow.UnboundedNavierStokesFlowSimulator3D(
        (n, n, n), 1.0, nu, real_t=real_t, time=1.0, with_forcing=True,
        with_free_stream_flow=with_free_stream, flow_density=rho,
    )
    sim.vorticity_field[...] = rng.random(sim.vorticity_field.shape).astype(real_t)
    sim.eul_grid_forcing_field[...] = rng.random(sim.eul_grid_forcing_field.shape).astype(real_t)
    w = sim.vorticity_field.copy()
    u = sim.velocity_field.copy()
    F = sim.eul_grid_forcing_field.copy()
    sim.time_step(dt, free_stream_velocity=U)
    dx = real_t(1.0 / n)
    stencils.update_vorticity_from_velocity_forcing(w, F, real_t(dt / (2 * dx * rho)))
    B = np.zeros_like(w)
    stencils.elementwise_cross_product(B, u, w)
    stencils.update_vorticity_from_velocity_forcing(w, B, real_t(dt / (2 * dx)))
    stencils.diffusion_timestep_euler_forward(w, np.zeros((n, n, n), dtype=real_t), real_t(nu * dt / dx / dx))
    pf = sim.position_field
    stencils.penalise_vector_field_boundary(w, 2, dx, pf[0], pf[1], pf[2])
    psi = np.zeros_like(w)
    poisson.UnboundedPoissonSolver3D(n, n, n, 1.0, real_t).vector_field_solve(psi, w)
    stencils.curl(u, psi, real_t(0.5 / dx))
    if with_free_stream:
        u += U.reshape(3, 1, 1, 1)
    assert sim.time == 1.0 + dt
    assert np.all(sim.eul_grid_forcing_field == 0)
    np.testing.assert_allclose(sim.vorticity_field, w, rtol=1e-7)
    np.testing.assert_allclose(sim.velocity_field, u, rtol=1e-7)


@pytest.mark.parametrize("n", [4, 8])
@pytest.mark.parametrize("real_t", [np.float32, np.float64])
def test_stable_timestep_closed_form(n, real_t):
    nu, cfl = 1e-2, 0.2
    sim = flow.UnboundedNavierStokesFlowSimulator3D((n, n, n), 1.0, nu, cfl=cfl, real_t=real_t)
    sim.velocity_field[...] = 2.0
    dx = real_t(1.0 / n)
    tol = 10 * np.finfo(real_t).eps
    ref = 0.5 * min(cfl * dx / (3 * 2.0 + tol), 0.9 * dx**2 / 6 / nu)
    assert sim.compute_stable_timestep(dt_prefac=0.5) == ref
