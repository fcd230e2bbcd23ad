"""Batched step (many simulators, one launch per stage, optional hipGraph replay): every
member's fields, virtual-boundary state, dt and time are bit-identical to stepping that
member alone with the reference driver loop (flow_past_sphere_case.py:191-197)."""

import numpy as np
import pytest

from helpers import PREC, to_host
from sopht_amd.utils import synthetic

pytestmark = pytest.mark.gpu

# (grid n, seed, sphere radius, Lagrangian nodes (0: no body), free stream)
SPECS = [
    (16, 1, 0.20, 200, (1.0, 0.0, 0.0)),
    (24, 2, 0.15, 300, (0.7, 0.1, 0.0)),
    (16, 3, 0.10, 0, (0.5, 0.0, 0.0)),
    (32, 4, 0.12, 400, (1.2, 0.0, 0.1)),
]


def make_member(spec, real_t):
    import torch

    from sopht_amd import simulator as S

    n, seed, radius, n_lag, U = spec
    flow = S.UnboundedNavierStokesFlowSimulator3D(
        grid_size=(n, n, n), x_range=1.0, kinematic_viscosity=1e-3, real_t=real_t,
        with_forcing=True, with_free_stream_flow=True,
    )
    w0 = synthetic.random_phase_vorticity((n, n, n), 1.0, 4, seed, 10.0).astype(real_t)
    flow.vorticity_field.copy_(torch.from_numpy(w0))
    U = np.asarray(U)
    flow.recover_velocity(U)
    inter = None
    if n_lag:
        inter = S.RigidBodyFlowInteraction(
            S.RigidBody((0.5, 0.5, 0.5), radius), flow.eul_grid_forcing_field, flow.velocity_field,
            -5e4, -10.0, flow.dx, 3, S.FibonacciSphereForcingGrid, real_t=real_t, num_lag_nodes=n_lag,
        )
    return flow, inter, U


def step_alone(flow, inter, U, steps):
    dts = []
    for _ in range(steps):
        dt = flow.compute_stable_timestep()
        if inter is not None:
            inter.time_step(dt)
            inter()
        flow.time_step(dt, free_stream_velocity=U)
        dts.append(float(dt))
    return dts


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("precision", ["single", "double"])
def test_batch_matches_members_stepped_alone(dev, precision, use_graph):
    from sopht_amd.simulator import BatchedFlowSimulation

    real_t = PREC[precision]
    steps = 4
    alone = [make_member(s, real_t) for s in SPECS]
    dts = [step_alone(f, i, U, steps) for f, i, U in alone]
    batch = [make_member(s, real_t) for s in SPECS]
    sim = BatchedFlowSimulation(batch)
    sim.step(1, use_graph=use_graph)   # the first step runs eagerly
    sim.step(steps - 1, use_graph=use_graph)
    last_dt = sim.synchronize()
    for k, ((fa, ia, _), (fb, ib, _)) in enumerate(zip(alone, batch)):
        assert fb.time == fa.time, k
        assert last_dt[k] == dts[k][-1], k
        for name in ("vorticity_field", "velocity_field", "stream_func_field", "eul_grid_forcing_field"):
            a, b = to_host(getattr(fa, name)), to_host(getattr(fb, name))
            assert np.array_equal(a, b), (k, name, float(np.abs(a - b).max()))
        if ia is not None:
            for name in ("lag_grid_position_mismatch_field", "lag_grid_velocity_mismatch_field",
                         "lag_grid_forcing_field", "lag_grid_flow_velocity_field"):
                assert np.array_equal(to_host(getattr(ia, name)), to_host(getattr(ib, name))), (k, name)
        # the batch wrote the member's velocity through raw pointers: dt must come from the
        # velocity it left, not from maxima cached before the batch
        assert fb.compute_stable_timestep() == fa.compute_stable_timestep(), k
    sim.close()


def test_batch_graph_replays_match_eager(dev):
    """Replays of the captured step, with eager host work and synchronisations in between,
    against the same batch stepped eagerly."""
    from sopht_amd.simulator import BatchedFlowSimulation

    real_t = np.float32
    eager = [make_member(s, real_t) for s in SPECS[:2]]
    graph = [make_member(s, real_t) for s in SPECS[:2]]
    se = BatchedFlowSimulation(eager)
    sg = BatchedFlowSimulation(graph)
    se.step(1, use_graph=False)
    sg.step(1, use_graph=True)
    for _ in range(3):
        se.step(2, use_graph=False)
        sg.step(2, use_graph=True)
        se.synchronize()
        sg.synchronize()
        for (fa, _, _), (fb, _, _) in zip(eager, graph):
            assert np.array_equal(to_host(fa.vorticity_field), to_host(fb.vorticity_field))
            assert fa.time == fb.time
