"""Parity at the workloads' own sizes.

* 64^3 FP64 coupled sphere (random-phase vorticity |k| <= 6 seed 42, max|w| = 10, nu = 1e-3,
  1,200-point Fibonacci sphere r = 0.1, U = (1, 0, 0), CFL 0.1), 50 steps against the
  oracle on the oracle's dt: relative L2 <= 1e-10 after step 1, and the drift of every step
  reported and bounded.
* 192^3 and 256^3 coupled sphere (|k| <= 8, nu = 5e-4, 3,000 points r = 0.08), one step
  against the FP64 oracle: FP64 relative L2 and Linf <= 1e-10 for the spread forcing, the
  vorticity, the velocity and the stream function; FP32 against the same oracle per field.
* 256^3: 200 coupled steps in FP32 against the same 200 steps in FP64 (the FP64 run's dt),
  the vorticity and velocity drift reported every 20 steps.
* The batched step with one member each of 48^3, 64^3, 96^3 and 128^3 (the batched
  workload's member sizes) bit-identical to stepping each member alone.
* The adjoint workload at 128^3 (200 steps, heaving ellipsoid, 2,000 points): the FP32
  gradient against the FP64 gradient on the same dt sequence.

The reference's own pattern for a composite step is ``test_navier_stokes_flow_simulators.py:163-308``;
the step under test is ``navier_stokes_flow_simulators.py:487-498``.
"""

import numpy as np
import pytest

from helpers import rel_l2, rel_linf, to_dev, to_host
from test_gpu_flow import build_pair, run_steps

pytestmark = pytest.mark.gpu


def sps():
    from sopht_amd import simulator

    return simulator


def test_c1_coupled_sphere_64_fp64_50_steps(dev):
    g, gi, o, oi, U = build_pair(64, np.float64, 1e-3, K=6, radius=0.1, n_lag=1200)
    errs = []
    for step in range(50):
        dt_g, dt_o = run_steps(g, gi, o, oi, U, 1)
        assert abs(dt_g - dt_o) <= 1e-10 * dt_o, step
        errs.append((rel_l2(to_host(g.vorticity_field), o.vorticity_field),
                     rel_l2(to_host(g.velocity_field), o.velocity_field)))
        if step == 0:
            assert max(errs[0]) <= 1e-10, errs[0]
    w = np.array(errs)
    print("64^3 FP64 drift per step (vorticity, velocity rel-L2):")
    for s in range(0, 50, 5):
        print(f"  step {s + 1:2d}: {w[s, 0]:.3e} {w[s, 1]:.3e}")
    print(f"  step 50: {w[-1, 0]:.3e} {w[-1, 1]:.3e}")
    assert w.max() <= 1e-9, w.max()
    ef = rel_l2(to_host(gi.lag_grid_forcing_field), oi.lag_grid_forcing_field)
    assert ef <= 1e-9, ef


def coupled_sphere_c2(n, real_t, w0):
    import torch

    S = sps()
    flow = S.UnboundedNavierStokesFlowSimulator3D(
        grid_size=(n, n, n), x_range=1.0, kinematic_viscosity=5e-4, cfl=0.1, real_t=real_t,
        with_forcing=True, with_free_stream_flow=True)
    flow.vorticity_field.copy_(torch.from_numpy(w0.astype(real_t)))
    U = np.array([1.0, 0.0, 0.0])
    flow.recover_velocity(U)
    inter = S.RigidBodyFlowInteraction(
        S.RigidBody((0.5, 0.5, 0.5), 0.08), flow.eul_grid_forcing_field, flow.velocity_field,
        -5e4, -10.0, flow.dx, 3, S.FibonacciSphereForcingGrid, real_t=real_t, num_lag_nodes=3000)
    return flow, inter, U


@pytest.mark.parametrize("n", [192, 256])
def test_c2_one_step_against_oracle(dev, n):
    import torch

    from oracle import flow as OF
    from oracle import poisson as OP
    from sopht_amd.utils import synthetic

    OP.WORKERS = 16
    S = sps()
    w0 = synthetic.random_phase_vorticity((n, n, n), 1.0, 8, 42, 10.0)
    U = np.array([1.0, 0.0, 0.0])
    o = OF.UnboundedNavierStokesFlowSimulator3D((n, n, n), 1.0, 5e-4, cfl=0.1, real_t=np.float64,
                                                with_forcing=True, with_free_stream_flow=True)
    o.vorticity_field[...] = w0
    o.recover_velocity(U)
    oi = OF.ImmersedBodyFlowInteraction(
        o, S.FibonacciSphereForcingGrid(3, S.RigidBody((0.5, 0.5, 0.5), 0.08), 3000), -5e4, -10.0, np.float64)
    dt = o.compute_stable_timestep()
    oi.time_step(dt)
    oi()
    F_o = o.eul_grid_forcing_field.copy()
    o.time_step(dt, free_stream_velocity=U)
    ref = {"forcing": F_o, "vorticity": o.vorticity_field, "velocity": o.velocity_field,
           "stream_function": o.stream_func_field}
    report = {}
    for real_t in (np.float64, np.float32):
        flow, inter, U = coupled_sphere_c2(n, real_t, w0)
        dt_g = flow.compute_stable_timestep()
        assert abs(dt_g - dt) <= (1e-12 if real_t == np.float64 else 1e-5) * dt
        inter.time_step(dt)
        inter()
        got = {"forcing": to_host(flow.eul_grid_forcing_field)}
        flow.time_step(dt, free_stream_velocity=U)
        got.update(vorticity=to_host(flow.vorticity_field), velocity=to_host(flow.velocity_field),
                   stream_function=to_host(flow.stream_func_field))
        for name, b in ref.items():
            report[(real_t.__name__, name)] = (rel_l2(got[name], b), rel_linf(got[name], b))
        del flow, inter, got
        torch.cuda.empty_cache()
    print(f"{n}^3 one coupled step against the FP64 oracle (rel-L2, rel-Linf):")
    for (prec, name), (e2, ei) in report.items():
        print(f"  {prec:8s} {name:16s} {e2:.3e} {ei:.3e}")
    for name in ref:
        e2, ei = report[("float64", name)]
        assert e2 <= 1e-10 and ei <= 1e-10, (name, e2, ei)
    # FP32: the vorticity bound of the workload's target, the others reported and bounded
    assert report[("float32", "vorticity")][0] <= 1e-5
    for name in ref:
        assert report[("float32", name)][0] <= 1e-4, (name, report[("float32", name)])


def test_c2_fp32_drift_200_steps_256(dev):
    import torch

    from sopht_amd.utils import synthetic

    n = 256
    w0 = synthetic.random_phase_vorticity((n, n, n), 1.0, 8, 42, 10.0)
    sims = [coupled_sphere_c2(n, rt, w0) for rt in (np.float64, np.float32)]
    del w0
    (f64, i64, U), (f32, i32, _) = sims
    drift = []
    for step in range(1, 201):
        dt = f64.compute_stable_timestep()
        for flow, inter, _ in sims:
            inter.time_step(dt)
            inter()
            flow.time_step(dt, free_stream_velocity=U)
        if step % 20 == 0:
            row = [step]
            for name in ("vorticity_field", "velocity_field"):
                a, b = getattr(f32, name).double(), getattr(f64, name)
                row.append((torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)).item())
            drift.append(row)
    i32.check_support_in_domain()
    print("256^3 FP32 vs FP64, 200 coupled steps (step, vorticity rel-L2, velocity rel-L2):")
    for s, ew, eu in drift:
        print(f"  {s:3d} {ew:.3e} {eu:.3e}")
    assert all(np.isfinite(r[1]) and np.isfinite(r[2]) for r in drift)
    assert drift[0][1] <= 1e-4 and drift[-1][1] <= 1e-2, drift


def test_c4_member_sizes_batched_bit_identical(dev):
    import torch

    from sopht_amd.simulator import BatchedFlowSimulation
    from test_gpu_batch import step_alone

    S = sps()
    from sopht_amd.utils import synthetic

    specs = [(48, 1, 0.05, (0.5, 0.0, 0.0)), (64, 2, 0.08, (1.0, 0.0, 0.0)),
             (96, 3, 0.10, (1.2, 0.0, 0.0)), (128, 4, 0.12, (1.5, 0.0, 0.0))]

    def member(spec):
        n, seed, radius, U = spec
        flow = S.UnboundedNavierStokesFlowSimulator3D(
            grid_size=(n, n, n), x_range=1.0, kinematic_viscosity=1e-3, real_t=np.float32,
            with_forcing=True, with_free_stream_flow=True)
        w0 = synthetic.random_phase_vorticity((n, n, n), 1.0, 6, seed, 10.0).astype(np.float32)
        flow.vorticity_field.copy_(torch.from_numpy(w0))
        U = np.asarray(U)
        flow.recover_velocity(U)
        inter = S.RigidBodyFlowInteraction(
            S.RigidBody((0.5, 0.5, 0.5), radius), flow.eul_grid_forcing_field, flow.velocity_field,
            -5e4, -10.0, flow.dx, 3, S.FibonacciSphereForcingGrid, real_t=np.float32, num_lag_nodes=1000)
        return flow, inter, U

    steps = 3
    alone = [member(s) for s in specs]
    dts = [step_alone(f, i, U, steps) for f, i, U in alone]
    batch = [member(s) for s in specs]
    sim = BatchedFlowSimulation(batch)
    sched = sim.poisson_schedules()
    assert sorted(c["grid"][0] for c in sched) == [48, 64, 96, 128], sched
    assert all(c["components"] == 3 and c["passes"] == 3 and c["zpass_jobs"] in (1, 2, 3, 6) for c in sched), sched
    sim.step(1)
    sim.step(steps - 1)
    last = sim.synchronize()
    for k, ((fa, ia, _), (fb, ib, _)) in enumerate(zip(alone, batch)):
        assert last[k] == dts[k][-1], k
        assert fb.time == fa.time, k
        for name in ("vorticity_field", "velocity_field", "stream_func_field", "eul_grid_forcing_field"):
            assert torch.equal(getattr(fa, name), getattr(fb, name)), (specs[k][0], name)
        for name in ("lag_grid_position_mismatch_field", "lag_grid_forcing_field"):
            assert torch.equal(getattr(ia, name), getattr(ib, name)), (specs[k][0], name)
    sim.close()


def test_c5_fp32_gradient_against_fp64_128(dev):
    """The adjoint workload (128^3, 200 steps, ellipsoid (0.1, 0.05, 0.05), 2,000 points,
    heave period 0.25, 16 coefficients ~ N(0, 0.1), checkpoints every 20): the FP32 gradient
    of the time-averaged x-force against the FP64 gradient on the FP64 run's dt sequence."""
    import torch

    from test_gpu_adjoint import build_rollout

    shape = (128, 128, 128)
    ns = 200
    grads, Js = {}, {}
    dts = None
    for real_t in (np.float64, np.float32):
        ro, *_ = build_rollout(shape, real_t, 2000, (0.1, 0.05, 0.05), 1e-3, 0.25, ns, 20)
        theta = torch.tensor(np.random.default_rng(5).normal(0.0, 0.1, 16), device=dev, requires_grad=True)
        forces = sps().body_forces(theta, ro, dts)
        if dts is None:
            dts = ro.dts.copy()
        J = sps().time_averaged_force(forces, dts)
        J.backward()
        grads[real_t] = to_host(theta.grad).astype(np.float64)
        Js[real_t] = float(J.detach())
        del ro, forces, J
        torch.cuda.empty_cache()
    err = rel_l2(grads[np.float32], grads[np.float64])
    ej = abs(Js[np.float32] - Js[np.float64]) / abs(Js[np.float64])
    print(f"128^3 adjoint workload: FP32 vs FP64 gradient rel-L2 {err:.3e}, objective rel {ej:.3e}, "
          f"|g64| {np.linalg.norm(grads[np.float64]):.3e}")
    assert err <= 1e-2, err
    assert ej <= 1e-3, ej
