"""Immersed-body interactor glue on the device: the reference's interactor tests
(``tests/test_simulator/test_immersed_body/test_immersed_body_flow_interaction.py:37-126``,
``cosserat_rod/test_cosserat_rod_flow_interaction.py:10-35``), in 3D (this backend's
dimension) with a latitude-ring sphere and a straight rod, and the flexible-filament
driver loop (``flow_past_rod_case.py:231-250``: rod substeps with the flow forces applied,
one spreading per flow step) against the oracle in FP64."""

import logging

import numpy as np
import pytest

from helpers import rel_l2, to_dev, to_host
from oracle import flow as OF
from sopht_amd.utils import synthetic

pytestmark = pytest.mark.gpu

LOGGER = "sopht_amd.simulator.immersed_body"


def sps():
    from sopht_amd import simulator

    return simulator


def sphere_interactor(dev, num_forcing_points_along_equator=16, dx=None, n=16, real_t=np.float64):
    """A latitude-ring sphere (radius 0.1 at the centre of the unit box) bound to random
    Eulerian fields."""
    import torch

    S = sps()
    dx = 1.0 / n if dx is None else dx
    rng = np.random.default_rng(0)
    u = torch.as_tensor(rng.random((3, n, n, n)).astype(real_t), device=dev)
    F = torch.zeros_like(u)
    body = S.RigidBody((0.5, 0.5, 0.5), 0.1, velocity=(0.3, -0.2, 0.1), omega=(0.0, 0.0, 4.0))
    inter = S.RigidBodyFlowInteraction(
        body, F, u, 1.0, 1.0, dx, 3, S.SphereForcingGrid, real_t=real_t,
        num_forcing_points_along_equator=num_forcing_points_along_equator,
    )
    return inter, F, u, dx


@pytest.mark.parametrize("num_forcing_points_along_equator", [4, 16, 128])
def test_interactor_resolution_notice(dev, num_forcing_points_along_equator, caplog):
    """``:36-77``: a warning when the Lagrangian spacing exceeds 2 dx or is below dx/2, an
    info notice otherwise, with the reference's text."""
    with caplog.at_level(logging.INFO):
        inter, _, _, dx = sphere_interactor(dev, num_forcing_points_along_equator)
    ds = inter.forcing_grid.get_maximum_lagrangian_grid_spacing()
    head = f"\n==================================================\nFor SphereForcingGrid:\nEulerian grid spacing (dx): {dx}"
    tail = "\n=================================================="
    if ds > 2 * dx:
        body = (f"\nMax Lagrangian grid spacing: {ds} > 2 * dx"
                "\nThe Lagrangian grid of the body is too coarse relative to"
                "\nthe Eulerian grid of the flow, which can lead to unexpected"
                "\nconvergence. Please make the Lagrangian grid finer.")
        level = logging.WARNING
    elif ds < 0.5 * dx:
        body = (f"\nMax Lagrangian grid spacing: {ds} < 0.5 * dx"
                "\nThe Lagrangian grid of the body is too fine relative to"
                "\nthe Eulerian grid of the flow, which corresponds to redundant"
                "\nforcing points. Please make the Lagrangian grid coarser.")
        level = logging.WARNING
    else:
        body = "\nLagrangian grid is resolved almost the same\nas the Eulerian grid of the flow."
        level = logging.INFO
    assert (LOGGER, level, head + body + tail) in caplog.record_tuples


def test_interactor_resolution_cases_cover_all_branches():
    """The three equator counts above fall in the three branches (spacing 2 pi r / m, dx 1/16)."""
    dx = 1.0 / 16
    ds = [0.1 * 2 * np.pi / m for m in (4, 16, 128)]
    assert ds[0] > 2 * dx and 0.5 * dx <= ds[1] <= 2 * dx and ds[2] < 0.5 * dx


def test_interactor_call_method(dev):
    """``:80-101``: ``interactor()`` equals grid kinematics + ``compute_interaction_forcing``."""
    import torch

    inter, F, u, _ = sphere_interactor(dev)
    inter()
    ref = torch.zeros_like(F)
    ref_inter, _, _, _ = sphere_interactor(dev)
    g = ref_inter.forcing_grid
    g.compute_lag_grid_position_field()
    g.compute_lag_grid_velocity_field()
    ref_inter.compute_interaction_forcing(
        eul_grid_forcing_field=ref, eul_grid_velocity_field=u,
        lag_grid_position_field=g.position_field, lag_grid_velocity_field=g.velocity_field,
    )
    assert torch.equal(ref, F)
    assert float(F.abs().max()) > 0


def test_interactor_compute_flow_forces_and_torques(dev):
    """``:104-126``: forces and torques equal the manual sequence (Lagrangian forcing, then the
    grid's transfer), and match the oracle's Lagrangian forcing of the same body."""
    inter, F, u, dx = sphere_interactor(dev)
    inter.compute_flow_forces_and_torques()
    g = inter.forcing_grid
    ref_f, ref_t = np.zeros((3, 1)), np.zeros((3, 1))
    g.compute_lag_grid_position_field()
    g.compute_lag_grid_velocity_field()
    inter.compute_interaction_force_on_lag_grid(
        eul_grid_velocity_field=u, lag_grid_position_field=g.position_field, lag_grid_velocity_field=g.velocity_field)
    g.transfer_forcing_from_grid_to_body(body_flow_forces=ref_f, body_flow_torques=ref_t,
                                         lag_grid_forcing_field=inter.lag_grid_forcing_field)
    np.testing.assert_array_equal(ref_f, inter.body_flow_forces)
    np.testing.assert_array_equal(ref_t, inter.body_flow_torques)
    assert np.abs(ref_f).max() > 0 and np.abs(ref_t).max() > 0
    # against the oracle: same grid, same velocity field, FP64
    n = u.shape[-1]

    class _Flow:
        pass

    flow = _Flow()
    flow.dx, flow.velocity_field, flow.eul_grid_forcing_field = dx, to_host(u), to_host(F)
    o = OF.ImmersedBodyFlowInteraction(flow, g, 1.0, 1.0, np.float64)
    o.compute_flow_forces_and_torques()
    for a, b in ((inter.body_flow_forces, o.body_flow_forces), (inter.body_flow_torques, o.body_flow_torques)):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * np.abs(b).max())
    assert n == 16


def test_interactor_grid_deviation_error_l2_norm(dev):
    """``:129-137``."""
    inter, _, _, _ = sphere_interactor(dev)
    inter.lag_grid_position_mismatch_field[...] = 2.0
    np.testing.assert_allclose(inter.get_grid_deviation_error_l2_norm(), 2.0 * np.sqrt(3))


@pytest.mark.parametrize("n_elems", [8, 16])
def test_cosserat_rod_flow_interaction(dev, n_elems):
    """``test_cosserat_rod_flow_interaction.py:10-35`` (3D): zero nodal forces (3, n+1) and
    element torques (3, n), and the requested grid."""
    import torch

    S = sps()
    rod = S.StraightRod(n_elems, np.array([0.3, 0.5, 0.5]), np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), 0.3, 0.005)
    u = torch.zeros((3, 16, 16, 16), dtype=torch.float64, device=dev)
    inter = S.CosseratRodFlowInteraction(
        cosserat_rod=rod, eul_grid_forcing_field=torch.zeros_like(u), eul_grid_velocity_field=u,
        virtual_boundary_stiffness_coeff=1.0, virtual_boundary_damping_coeff=1.0, dx=1.0 / 16, grid_dim=3,
        real_t=np.float64, forcing_grid_cls=S.CosseratRodElementCentricForcingGrid,
    )
    np.testing.assert_array_equal(inter.body_flow_forces, np.zeros((3, n_elems + 1)))
    np.testing.assert_array_equal(inter.body_flow_torques, np.zeros((3, n_elems)))
    assert isinstance(inter.forcing_grid, S.CosseratRodElementCentricForcingGrid)


@pytest.mark.parametrize("grid_cls", ["CosseratRodNodalForcingGrid", "CosseratRodSurfaceForcingGrid"])
def test_rod_coupled_loop_fp64_matches_oracle(dev, grid_cls):
    """The flexible-filament driver loop on a heaving rod (32^3 FP64, 4 flow steps with
    their rod substeps): flow fields, interactor state and the rod's external loads against
    the oracle running the same loop with the same host rod and grid classes."""
    S = sps()
    n, real_t = 32, np.float64
    U = np.array([1.0, 0.0, 0.0])
    kw = dict(kinematic_viscosity=1e-3, real_t=real_t, with_forcing=True, with_free_stream_flow=True)
    g = S.UnboundedNavierStokesFlowSimulator3D(grid_size=(n, n, n), x_range=1.0, **kw)
    o = OF.UnboundedNavierStokesFlowSimulator3D((n, n, n), 1.0, **kw)
    w0 = synthetic.random_phase_vorticity((n, n, n), 1.0, 4, 11, 10.0)
    g.vorticity_field.copy_(to_dev(w0, dev))
    o.vorticity_field[...] = w0
    g.recover_velocity(U)
    o.recover_velocity(U)

    def rod_and_grid_kwargs():
        rod = S.StraightRod(16, np.array([0.3, 0.5, 0.5]), np.array([1.0, 0, 0]), np.array([0, 1.0, 0]), 0.3, 0.02)
        extra = {"surface_grid_density_for_largest_element": 6} if grid_cls == "CosseratRodSurfaceForcingGrid" else {}
        return rod, extra

    rod_g, extra = rod_and_grid_kwargs()
    rod_o, _ = rod_and_grid_kwargs()
    cls = getattr(S, grid_cls)
    gi = S.CosseratRodFlowInteraction(rod_g, g.eul_grid_forcing_field, g.velocity_field, -2e4, -1e2, g.dx, 3, cls,
                                      real_t=real_t, **extra)
    grid_o = cls(grid_dim=3, cosserat_rod=rod_o, **extra)
    oi = OF.ImmersedBodyFlowInteraction(o, grid_o, -2e4, -1e2, real_t,
                                        body_flow_forces=np.zeros((3, 17)), body_flow_torques=np.zeros((3, 16)))
    kin_g, kin_o = synthetic.HeavingRod(rod_g), synthetic.HeavingRod(rod_o)
    ff_g, ff_o = S.FlowForces(gi), S.FlowForces(oi)
    rod_dt = 0.01 * 0.3 / 16
    for step in range(4):
        dt_g = synthetic.rod_coupled_step(g, gi, ff_g, rod_g, kin_g, rod_dt, U)
        dt_o = synthetic.rod_coupled_step(o, oi, ff_o, rod_o, kin_o, rod_dt, U)
        assert abs(dt_g - dt_o) <= 1e-12 * dt_o, step
        for name in ("vorticity_field", "velocity_field"):
            e = rel_l2(to_host(getattr(g, name)), getattr(o, name))
            assert e <= 1e-10, (step, name, e)
        for a, b in ((rod_g.external_forces, rod_o.external_forces),
                     (rod_g.external_torques, rod_o.external_torques),
                     (to_host(gi.lag_grid_position_mismatch_field), oi.lag_grid_position_mismatch_field)):
            np.testing.assert_allclose(a, b, rtol=0, atol=1e-10 * np.abs(b).max(), err_msg=str(step))
    assert np.abs(rod_g.external_forces).max() > 0
    assert g.time == pytest.approx(o.time, rel=1e-14)
