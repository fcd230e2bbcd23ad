"""Checkpoint/restart of the coupled flow on the GPU (``sopht/utils/io.py``,
``sopht/utils/restart_sim.py:11-37``, the reference's ``tests/test_utils/test_restart.py``).

A run saves the flow (vorticity, velocity through ``EulerianFieldIO``) and the forcing grid
(positions, forcing, position and velocity mismatch through ``IO``) half way, then continues.
The velocity mismatch is state too: the next ``time_step`` integrates the position mismatch
with the mismatch of the previous forcing call (``VirtualBoundaryForcing.py:276-283``); the
reference's restart example does not save it, so its first resumed step integrates with zero. A fresh
simulator restarted from those files and stepped the same number of times must reproduce the
continued run bit for bit: every kernel of the step is deterministic and the files hold the
fields exactly, so the restart is exact, not approximate.
"""

import numpy as np
import pytest
import torch

from sopht_amd.utils import synthetic

pytestmark = pytest.mark.gpu


def make_case(n, real_t, filter_vorticity):
    from sopht_amd import simulator as S

    kw = dict(kinematic_viscosity=3e-3, real_t=real_t, with_forcing=True, with_free_stream_flow=True)
    if filter_vorticity:
        kw.update(filter_vorticity=True, filter_setting_dict={"type": "multiplicative", "order": 1})
    flow = S.UnboundedNavierStokesFlowSimulator3D(grid_size=(n, n, n), x_range=1.0, **kw)
    w0 = synthetic.random_phase_vorticity((n, n, n), 1.0, 6, 7, 10.0).astype(real_t)
    flow.vorticity_field.copy_(torch.from_numpy(w0).to(flow.vorticity_field.device))
    U = np.array([1.0, 0.0, 0.0])
    flow.recover_velocity(U)
    body = S.RigidBody((0.5, 0.5, 0.5), 0.12)
    inter = S.RigidBodyFlowInteraction(
        body, flow.eul_grid_forcing_field, flow.velocity_field, -5e4, -10.0, flow.dx, 3,
        S.FibonacciSphereForcingGrid, real_t=real_t, num_lag_nodes=1500)
    return flow, inter, U


def ios(flow, inter):
    import sopht_amd.utils as spu

    io = spu.EulerianFieldIO(position_field=flow.position_field,
                             eulerian_fields_dict={"vorticity": flow.vorticity_field, "velocity": flow.velocity_field})
    forcing_io = spu.IO(dim=3, real_dtype=flow.real_t)
    forcing_io.add_as_lagrangian_fields_for_io(
        lagrangian_grid=inter.forcing_grid.position_field, lagrangian_grid_name="sphere",
        lag_grid_forcing_field=inter.lag_grid_forcing_field,
        lag_grid_position_mismatch_field=inter.lag_grid_position_mismatch_field,
        lag_grid_velocity_mismatch_field=inter.lag_grid_velocity_mismatch_field)
    return io, forcing_io


def advance(flow, inter, U, steps):
    dts = []
    for _ in range(steps):
        dt = flow.compute_stable_timestep()
        inter.time_step(dt)
        inter()
        flow.time_step(dt, free_stream_velocity=U)
        dts.append(dt)
    return dts


@pytest.mark.parametrize("real_t", [np.float32, np.float64])
@pytest.mark.parametrize("filter_vorticity", [False, True])
def test_restart_reproduces_continued_run(tmp_path, real_t, filter_vorticity):
    import sopht_amd.utils as spu

    n, first, second = 32, 4, 5
    flow, inter, U = make_case(n, real_t, filter_vorticity)
    io, forcing_io = ios(flow, inter)
    advance(flow, inter, U, first)
    step = 100
    io.save(h5_file_name=str(tmp_path / f"sopht_{step:04d}.h5"), time=flow.time)
    forcing_io.save(h5_file_name=str(tmp_path / f"forcing_grid_{step:04d}.h5"), time=flow.time)
    dts_full = advance(flow, inter, U, second)
    torch.cuda.synchronize()

    fresh, fresh_inter, _ = make_case(n, real_t, filter_vorticity)
    fresh.vorticity_field.zero_()
    fresh.velocity_field.zero_()
    io2, forcing_io2 = ios(fresh, fresh_inter)
    fresh.time = spu.restart_simulation(None, io2, None, forcing_io2, str(tmp_path / "restart_data"),
                                        h5_dir=str(tmp_path))
    dts_restarted = advance(fresh, fresh_inter, U, second)
    torch.cuda.synchronize()

    assert dts_restarted == dts_full
    assert fresh.time == pytest.approx(flow.time, rel=0, abs=1e-15)
    for a, b in ((flow.vorticity_field, fresh.vorticity_field), (flow.velocity_field, fresh.velocity_field),
                 (inter.lag_grid_position_mismatch_field, fresh_inter.lag_grid_position_mismatch_field),
                 (inter.lag_grid_velocity_mismatch_field, fresh_inter.lag_grid_velocity_mismatch_field),
                 (inter.lag_grid_forcing_field, fresh_inter.lag_grid_forcing_field)):
        assert torch.equal(a, b)
    assert torch.count_nonzero(inter.lag_grid_position_mismatch_field) > 0
