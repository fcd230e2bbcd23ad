"""GPU parity of the whole coupled flow step (drop-in simulator + interactor) against the
oracle on the same seeded inputs.

* One step, reference integration-test pattern (``test_navier_stokes_flow_simulators.py:163-308``):
  random vorticity and forcing, n in {4, 8, 16}, free stream on/off, rho in {1, 0.1}.
* FP64: relative L2 and Linf <= 1e-10 per field after one step; multi-step trajectories
  with the immersed sphere report and bound the drift.
* FP32: error against the FP64 oracle per field, bounded.
"""

import numpy as np
import pytest

from helpers import PREC, rel_l2, rel_linf, to_dev, to_host, tol
from oracle import flow as OF
from sopht_amd.utils import synthetic

pytestmark = pytest.mark.gpu


def sps():
    from sopht_amd import simulator

    return simulator


@pytest.mark.parametrize("n", [4, 8, 16])
@pytest.mark.parametrize("with_free_stream", [True, False])
@pytest.mark.parametrize("rho", [1.0, 0.1])
def test_one_step_fp64_matches_oracle(dev, n, with_free_stream, rho, rng):
    real_t = np.float64
    kw = dict(kinematic_viscosity=1e-2, real_t=real_t, time=1.0, with_forcing=True,
              with_free_stream_flow=with_free_stream, flow_density=rho)
    o = OF.UnboundedNavierStokesFlowSimulator3D((n, n, n), 1.0, **kw)
    g = sps().UnboundedNavierStokesFlowSimulator3D(grid_size=(n, n, n), x_range=1.0, **kw)
    w0 = rng.random((3, n, n, n))
    u0 = rng.random((3, n, n, n))
    F0 = rng.random((3, n, n, n))
    o.vorticity_field[...] = w0
    o.velocity_field[...] = u0
    o.eul_grid_forcing_field[...] = F0
    g.vorticity_field.copy_(to_dev(w0, dev))
    g.velocity_field.copy_(to_dev(u0, dev))
    g.eul_grid_forcing_field.copy_(to_dev(F0, dev))
    U = np.array([3.0, 4.0, 5.0])
    o.time_step(dt=2.0 * 1e-3, free_stream_velocity=U)
    g.time_step(dt=2.0 * 1e-3, free_stream_velocity=U)
    assert g.time == o.time
    assert np.all(to_host(g.eul_grid_forcing_field) == 0)
    for name in ("vorticity_field", "velocity_field", "stream_func_field"):
        a, b = to_host(getattr(g, name)), getattr(o, name)
        assert rel_l2(a, b) <= 1e-10, (name, rel_l2(a, b))
        assert rel_linf(a, b) <= 1e-10, (name, rel_linf(a, b))


@pytest.mark.parametrize("shape", [(8, 12, 6), (12, 8, 16), (6, 6, 6), (24, 16, 12)])
@pytest.mark.parametrize("precision", ["single", "double"])
def test_one_step_non_cubic_matches_oracle(dev, shape, precision, rng):
    """Non-cubic grids and row lengths that are not a multiple of the 16-byte vector (the
    sweeps' scalar path) against the FP64 oracle: 1e-10 in FP64; in FP32 the oracle runs on
    the FP32-rounded inputs and the bound is 2e-5 (relative L2) per field."""
    real_t = PREC[precision]
    kw = dict(kinematic_viscosity=1e-2, time=0.0, with_forcing=True, with_free_stream_flow=True)
    o = OF.UnboundedNavierStokesFlowSimulator3D(shape, 1.0, real_t=np.float64, **kw)
    g = sps().UnboundedNavierStokesFlowSimulator3D(grid_size=shape, x_range=1.0, real_t=real_t, **kw)
    w0, u0, F0 = (rng.random((3,) + shape).astype(real_t) for _ in range(3))
    o.vorticity_field[...] = w0
    o.velocity_field[...] = u0
    o.eul_grid_forcing_field[...] = F0
    g.vorticity_field.copy_(to_dev(w0, dev))
    g.velocity_field.copy_(to_dev(u0, dev))
    g.eul_grid_forcing_field.copy_(to_dev(F0, dev))
    U = np.array([1.0, 0.5, -0.25])
    o.time_step(dt=1e-3, free_stream_velocity=U)
    g.time_step(dt=1e-3, free_stream_velocity=U)
    bound = 1e-10 if real_t == np.float64 else 2e-5
    assert np.all(to_host(g.eul_grid_forcing_field) == 0)
    for name in ("vorticity_field", "velocity_field", "stream_func_field"):
        a, b = to_host(getattr(g, name)), getattr(o, name)
        assert rel_l2(a, b) <= bound, (name, rel_l2(a, b))


@pytest.mark.parametrize("shape", [(24, 24, 48), (12, 48, 96), (64, 12, 128), (48, 16, 8), (4, 4, 4)])
@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("chunk", [16, 32])
@pytest.mark.parametrize("cols", [64, 128])
def test_vector_s1_sweep_bit_identical_to_scalar(dev, shape, precision, chunk, cols, rng, monkeypatch):
    """The S1 sweep with 16-byte accesses (rows a multiple of 16 bytes; forced here with
    ``SOPHT_S1_CHUNK``, as these grids are below the size that selects it) against the scalar
    sweep (``SOPHT_S1_SCALAR=1``): the same bits in every field after one step, with partial
    x, y and z tiles, forcing nonzero only in a sub-box (the F-zeroing flags per tile, read by
    SC's tiles of 16 planes), and the out-of-place step's omega1 output. FP32 runs both tile
    widths (``SOPHT_S1_COLS``: 64 cells by 16 rows, 128 by 8); FP64 has the 64-cell tile."""
    import torch

    real_t = PREC[precision]
    kw = dict(kinematic_viscosity=1e-2, time=0.0, with_forcing=True, with_free_stream_flow=True)
    sims = []
    for scalar in (False, True):
        if scalar:
            monkeypatch.delenv("SOPHT_S1_CHUNK")
            monkeypatch.setenv("SOPHT_S1_SCALAR", "1")
        else:
            monkeypatch.setenv("SOPHT_S1_CHUNK", str(chunk))
            monkeypatch.setenv("SOPHT_S1_COLS", str(cols))
        sims.append(sps().UnboundedNavierStokesFlowSimulator3D(grid_size=shape, x_range=1.0,
                                                               real_t=real_t, **kw))
    monkeypatch.delenv("SOPHT_S1_SCALAR")
    monkeypatch.delenv("SOPHT_S1_COLS")
    from sopht_amd import _lib as L

    want_cols = cols if precision == "single" else 64
    assert L.schedule("sopht_ns3d_s1_schedule", sims[0]._plan) == (1, chunk, want_cols)
    assert L.schedule("sopht_ns3d_s1_schedule", sims[1]._plan) == (0, 16, 64)
    w0, u0 = (rng.standard_normal((3,) + shape).astype(real_t) for _ in range(2))
    F0 = np.zeros((3,) + shape, dtype=real_t)
    nz, ny, nx = shape
    F0[:, nz // 4: nz // 2 + 1, ny // 3: ny // 2 + 2, nx // 5: nx // 2 + 1] = rng.standard_normal(
        (3, nz // 2 + 1 - nz // 4, ny // 2 + 2 - ny // 3, nx // 2 + 1 - nx // 5)).astype(real_t)
    U = np.array([1.0, 0.5, -0.25])
    for g in sims:
        g.vorticity_field.copy_(to_dev(w0, dev))
        g.velocity_field.copy_(to_dev(u0, dev))
        g.eul_grid_forcing_field.copy_(to_dev(F0, dev))
        g.time_step(dt=1e-2, free_stream_velocity=U)
    a, b = sims
    for name in ("vorticity_field", "velocity_field", "stream_func_field", "eul_grid_forcing_field"):
        assert torch.equal(getattr(a, name), getattr(b, name)), name
    assert np.all(to_host(a.eul_grid_forcing_field) == 0)
    outs = []
    for g in sims:
        win, uin, F = (to_dev(x, dev) for x in (w0, u0, F0))
        wo, uo, w1 = torch.empty_like(win), torch.empty_like(uin), torch.empty_like(win)
        g.trajectory_step(win, uin, F, wo, uo, 1e-2, U, vorticity1_out=w1)
        outs.append((wo, uo, w1))
    for x, y in zip(*outs):
        assert torch.equal(x, y)


@pytest.mark.parametrize("precision", ["single", "double"])
def test_stable_timestep(dev, precision):
    real_t = PREC[precision]
    n, nu, cfl = 8, 1e-2, 0.2
    g = sps().UnboundedNavierStokesFlowSimulator3D((n, n, n), 1.0, nu, cfl=cfl, real_t=real_t)
    o = OF.UnboundedNavierStokesFlowSimulator3D((n, n, n), 1.0, nu, cfl=cfl, real_t=real_t)
    g.velocity_field.fill_(2.0)
    o.velocity_field[...] = 2.0
    assert g.compute_stable_timestep(dt_prefac=0.5) == o.compute_stable_timestep(dt_prefac=0.5)
    rng = np.random.default_rng(3)
    u = rng.standard_normal((3, n, n, n)).astype(real_t)
    g.velocity_field.copy_(to_dev(u, dev))
    o.velocity_field[...] = u
    assert g.compute_stable_timestep() == o.compute_stable_timestep()


@pytest.mark.parametrize("precision", ["single", "double"])
def test_stable_timestep_after_drop_in_write(dev, precision):
    """A drop-in kernel that writes the velocity field after a step (through raw pointers)
    must not leave dt on the maxima the step's velocity sweep kept."""
    from sopht_amd.numeric import eulerian_grid_ops as spne

    real_t = PREC[precision]
    n = 16
    g = sps().UnboundedNavierStokesFlowSimulator3D((n, n, n), 1.0, 1e-3, real_t=real_t,
                                                 with_free_stream_flow=True)
    w0 = synthetic.random_phase_vorticity((n, n, n), 1.0, 4, 5, 10.0).astype(real_t)
    g.vorticity_field.copy_(to_dev(w0, dev))
    g.time_step(1e-3, free_stream_velocity=np.array([1.0, 0.0, 0.0]))
    before = g.compute_stable_timestep()
    add = spne.gen_add_fixed_val_pyst_kernel_3d(real_t=real_t, field_type="vector")
    add(sum_field=g.velocity_field, vector_field=g.velocity_field, fixed_vals=[3.0, 0.0, 0.0])
    vmax = real_t(g.velocity_field.abs().sum(dim=0).max().item())
    assert g.max_abs_velocity_sum() == vmax
    assert g.compute_stable_timestep() == g._stable_timestep_from_max(vmax)
    assert g.compute_stable_timestep() < before


def test_vorticity_divergence_norm(dev, rng):
    n = 16
    w = rng.random((3, n, n, n))
    g = sps().UnboundedNavierStokesFlowSimulator3D((n, n, n), 1.0, 1e-2, real_t=np.float64)
    o = OF.UnboundedNavierStokesFlowSimulator3D((n, n, n), 1.0, 1e-2, real_t=np.float64)
    g.vorticity_field.copy_(to_dev(w, dev))
    o.vorticity_field[...] = w
    assert g.get_vorticity_divergence_l2_norm() == pytest.approx(o.get_vorticity_divergence_l2_norm(), rel=1e-12)


def build_pair(n, real_t, nu, steps_seed=42, K=6, radius=0.1, n_lag=1200, k=-5e4, c=-10.0):
    """GPU simulator + interactor and the oracle twin on the same seeded inputs
    (random-phase vorticity, Fibonacci sphere at the centre)."""
    S = sps()
    kw = dict(kinematic_viscosity=nu, real_t=real_t, with_forcing=True, with_free_stream_flow=True)
    g = S.UnboundedNavierStokesFlowSimulator3D(grid_size=(n, n, n), x_range=1.0, **kw)
    o = OF.UnboundedNavierStokesFlowSimulator3D((n, n, n), 1.0, **kw)
    w0 = synthetic.random_phase_vorticity((n, n, n), 1.0, K, steps_seed, 10.0).astype(real_t)
    g.vorticity_field.copy_(to_dev(w0, dev_of(g)))
    o.vorticity_field[...] = w0
    U = np.array([1.0, 0.0, 0.0])
    g.recover_velocity(U)
    o.recover_velocity(U)
    body = S.RigidBody((0.5, 0.5, 0.5), radius)
    gi = S.RigidBodyFlowInteraction(
        body, g.eul_grid_forcing_field, g.velocity_field, k, c, g.dx, 3,
        S.FibonacciSphereForcingGrid, real_t=real_t, num_lag_nodes=n_lag)
    grid = S.FibonacciSphereForcingGrid(3, body, n_lag)
    oi = OF.ImmersedBodyFlowInteraction(o, grid, k, c, real_t)
    return g, gi, o, oi, U


def dev_of(sim):
    return sim.vorticity_field.device


def run_steps(g, gi, o, oi, U, steps):
    for _ in range(steps):
        dt_o = o.compute_stable_timestep()
        dt_g = g.compute_stable_timestep()
        # the trajectories share the oracle's dt so drift measures the step alone
        gi.time_step(dt_o)
        gi()
        g.time_step(dt_o, free_stream_velocity=U)
        oi.time_step(dt_o)
        oi()
        o.time_step(dt_o, free_stream_velocity=U)
    return dt_g, dt_o


def test_coupled_trajectory_fp64_32cubed(dev):
    g, gi, o, oi, U = build_pair(32, np.float64, 1e-3, n_lag=300, radius=0.15)
    errs = []
    for _ in range(5):
        dt_g, dt_o = run_steps(g, gi, o, oi, U, 2)
        errs.append(rel_l2(to_host(g.vorticity_field), o.vorticity_field))
        assert abs(dt_g - dt_o) <= 1e-10 * dt_o
    assert max(errs) <= 1e-10, errs
    assert rel_l2(to_host(g.velocity_field), o.velocity_field) <= 1e-10
    assert rel_l2(to_host(gi.lag_grid_forcing_field), oi.lag_grid_forcing_field) <= 1e-10
    print("fp64 32^3 vorticity drift per 2 steps:", errs)


def test_coupled_step_fp32_vs_fp64_oracle(dev):
    g, gi, _, _, U = build_pair(32, np.float32, 1e-3, n_lag=300, radius=0.15)
    _, _, o, oi, _ = build_pair(32, np.float64, 1e-3, n_lag=300, radius=0.15)
    run_steps(g, gi, o, oi, U, 3)
    ew = rel_l2(to_host(g.vorticity_field), o.vorticity_field)
    eu = rel_l2(to_host(g.velocity_field), o.velocity_field)
    print(f"fp32 vs fp64 oracle after 3 steps: vorticity rel-L2 {ew:.3e}, velocity rel-L2 {eu:.3e}")
    assert ew < 1e-5 and eu < 1e-5


def test_benchmark_workload_fp32_vs_fp64_at_256(dev):
    """The benchmarked workload at its full size (256^3, random-phase vorticity |k| <= 8 seed 42
    scaled to max|w| = 10, nu = 5e-4, sphere r = 0.08 with 3000 Fibonacci points, U = (1, 0, 0)):
    three coupled steps in FP32 against the same steps in FP64, which match the oracle to 1e-10
    on every grid the oracle covers; both take the FP64 run's dt. Vorticity and velocity
    relative L2 <= 1e-5."""
    import torch

    S = sps()
    n = 256
    w0 = synthetic.random_phase_vorticity((n, n, n), 1.0, 8, 42, 10.0)
    U = np.array([1.0, 0.0, 0.0])
    sims = []
    for real_t in (np.float64, np.float32):
        flow = S.UnboundedNavierStokesFlowSimulator3D(
            grid_size=(n, n, n), x_range=1.0, kinematic_viscosity=5e-4, cfl=0.1, real_t=real_t,
            with_forcing=True, with_free_stream_flow=True)
        flow.vorticity_field.copy_(torch.from_numpy(w0.astype(real_t)))
        flow.recover_velocity(U)
        inter = S.RigidBodyFlowInteraction(
            S.RigidBody((0.5, 0.5, 0.5), 0.08), flow.eul_grid_forcing_field, flow.velocity_field,
            -5e4, -10.0, flow.dx, 3, S.FibonacciSphereForcingGrid, real_t=real_t, num_lag_nodes=3000)
        sims.append((flow, inter))
    del w0
    (f64, i64), (f32, i32) = sims
    for _ in range(3):
        dt = f64.compute_stable_timestep()
        for flow, inter in sims:
            inter.time_step(dt)
            inter()
            flow.time_step(dt, free_stream_velocity=U)
    i32.check_support_in_domain()
    for name in ("vorticity_field", "velocity_field"):
        a = getattr(f32, name).double()
        b = getattr(f64, name)
        err = (torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)).item()
        print(f"256^3 fp32 vs fp64 after 3 steps: {name} rel-L2 {err:.3e}")
        assert err <= 1e-5, (name, err)


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("shape", [(16, 16, 16), (8, 12, 6)])
def test_fused_velocity_max_matches_full_reduction(dev, precision, shape, rng):
    """The CFL maximum from the velocity sweep's per-workgroup maxima equals the full
    reduction over the field bit for bit; once the field changes through torch the full
    reduction is used (passive_transport_flow_simulators.py:139-155)."""
    import torch

    from sopht_amd import _lib as L

    real_t = PREC[precision]
    g = sps().UnboundedNavierStokesFlowSimulator3D(
        grid_size=shape, x_range=1.0, kinematic_viscosity=1e-2, real_t=real_t, with_forcing=True,
        with_free_stream_flow=True)
    g.vorticity_field.copy_(to_dev(rng.standard_normal((3,) + shape).astype(real_t), dev))
    g.eul_grid_forcing_field.copy_(to_dev(rng.standard_normal((3,) + shape).astype(real_t), dev))
    U = np.array([0.3, -0.2, 0.1])
    g.recover_velocity(U)
    g.time_step(dt=1e-3, free_stream_velocity=U)

    def full():
        out = torch.zeros((), dtype=g.velocity_field.dtype, device=dev)
        L.call("sopht_ns3d_max_abs_velocity_sum", g._plan, L.ptr(g.velocity_field), L.ptr(out),
               L.stream_ptr())
        return out.item()

    ref = np.max(np.sum(np.abs(to_host(g.velocity_field)), axis=0))
    assert g.max_abs_velocity_sum() == real_t(full()) == ref
    g.velocity_field.mul_(2.0)  # a change the step did not make
    assert g.max_abs_velocity_sum() == real_t(full()) == real_t(2.0) * ref


@pytest.mark.parametrize("filter_type", ["multiplicative", "convolution"])
@pytest.mark.parametrize("filter_order", [1, 2, 5, 9])
@pytest.mark.parametrize("precision", ["single", "double"])
def test_one_step_with_vorticity_filter(dev, filter_type, filter_order, precision, rng):
    """filter_vorticity=True (navier_stokes_flow_simulators.py:274-292,396-408,473): the fused
    step with the Laplacian filter between diffusion and penalisation against the oracle."""
    real_t = PREC[precision]
    shape = (16, 12, 16)
    setting = {"order": filter_order, "type": filter_type}
    _filtered_step_against_oracle(dev, shape, setting, real_t, rng)


@pytest.mark.parametrize("shape", [(24, 12, 96), (12, 8, 384), (6, 48, 64)])
@pytest.mark.parametrize("precision", ["single", "double"])
def test_filtered_step_order5_convolution_tiles(dev, shape, precision, rng):
    """The rod example's filter (order 5, convolution: flow_past_rod_case.py:103-104) across
    row and line-block edges of the fused per-axis filter kernels (rows longer than one and
    several wavefronts, more x than one line block), and the multiplicative order 2 default."""
    for setting in ({"order": 5, "type": "convolution"}, {"order": 2, "type": "multiplicative"}):
        _filtered_step_against_oracle(dev, shape, setting, PREC[precision], rng)


def _filtered_step_against_oracle(dev, shape, setting, real_t, rng):
    kw = dict(kinematic_viscosity=1e-2, time=0.0, with_forcing=True, with_free_stream_flow=True)
    o = OF.UnboundedNavierStokesFlowSimulator3D(shape, 1.0, real_t=np.float64, filter_setting_dict=setting, **kw)
    g = sps().UnboundedNavierStokesFlowSimulator3D(grid_size=shape, x_range=1.0, real_t=real_t,
                                                  filter_vorticity=True, filter_setting_dict=setting, **kw)
    w0, u0, F0 = (rng.random((3,) + shape).astype(real_t) for _ in range(3))
    o.vorticity_field[...] = w0
    o.velocity_field[...] = u0
    o.eul_grid_forcing_field[...] = F0
    g.vorticity_field.copy_(to_dev(w0, dev))
    g.velocity_field.copy_(to_dev(u0, dev))
    g.eul_grid_forcing_field.copy_(to_dev(F0, dev))
    U = np.array([1.0, 0.5, -0.25])
    o.time_step(dt=1e-3, free_stream_velocity=U)
    g.time_step(dt=1e-3, free_stream_velocity=U)
    bound = 1e-10 if real_t == np.float64 else 2e-5
    for name in ("vorticity_field", "velocity_field", "stream_func_field"):
        a, b = to_host(getattr(g, name)), getattr(o, name)
        assert rel_l2(a, b) <= bound, (name, rel_l2(a, b))


@pytest.mark.parametrize("grid_size_x", [4, 8, 16])
@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("field_type", ["scalar", "vector"])
def test_passive_transport_time_step(dev, grid_size_x, precision, field_type, rng):
    """PassiveTransportFlowSimulator.time_step against the manual ENO3 advection + diffusion
    sequence of the oracle (test_passive_transport_flow_simulators.py:8-88)."""
    from oracle import stencils as OS

    real_t = PREC[precision]
    nu, dt, init_time = 1e-2, 2.0, 1.0
    grid = (grid_size_x,) * 3
    dx = real_t(1.0 / grid_size_x)
    sim = sps().PassiveTransportFlowSimulator(
        kinematic_viscosity=nu, grid_dim=3, grid_size=grid, x_range=1.0, real_t=real_t,
        time=init_time, field_type=field_type)
    u = rng.random(tuple(sim.velocity_field.shape)).astype(real_t)
    f = rng.random(tuple(sim.primary_field.shape)).astype(real_t)
    sim.velocity_field.copy_(to_dev(u, dev))
    sim.primary_field.copy_(to_dev(f, dev))
    sim.time_step(dt=dt)
    ref = f.copy()
    buf = np.zeros(grid, dtype=real_t)
    OS.advection_timestep_euler_forward_conservative_eno3(ref, buf, u, real_t(dt / dx))
    OS.diffusion_timestep_euler_forward(ref, buf, real_t(nu * dt / dx / dx))
    assert sim.time == init_time + dt
    # dt/dx up to 32 and nu dt/dx^2 up to 5.1 amplify rounding (values grow to ~50): the
    # reference's tolerance 1e3 eps, relative and absolute
    np.testing.assert_allclose(to_host(sim.primary_field), ref, rtol=tol(real_t), atol=tol(real_t))


@pytest.mark.parametrize("shape", [(20, 37, 70), (33, 9, 130)])
@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("field_type", ["scalar", "vector"])
def test_passive_transport_fused_sweep_tiles(dev, shape, precision, field_type, rng):
    """The fused transport sweep across tile, z-chunk and grid edges (several x and y tiles,
    partial last tiles, z chunks) over three steps, against the oracle's advection step then
    diffusion step per component."""
    from oracle import stencils as OS

    real_t = PREC[precision]
    nu, dt = 1e-3, 2e-3
    sim = sps().PassiveTransportFlowSimulator(
        kinematic_viscosity=nu, grid_dim=3, grid_size=shape, x_range=1.0, real_t=real_t, field_type=field_type)
    dx = sim.dx
    u = (rng.random(tuple(sim.velocity_field.shape)) - 0.5).astype(real_t)
    f = rng.random(tuple(sim.primary_field.shape)).astype(real_t)
    sim.velocity_field.copy_(to_dev(u, dev))
    sim.primary_field.copy_(to_dev(f, dev))
    ref = f.copy()
    buf = np.zeros(shape, dtype=real_t)
    for _ in range(3):
        sim.time_step(dt=dt)
        OS.advection_timestep_euler_forward_conservative_eno3(ref, buf, u, real_t(dt / dx))
        OS.diffusion_timestep_euler_forward(ref, buf, real_t(nu * dt / dx / dx))
    np.testing.assert_allclose(to_host(sim.primary_field), ref, rtol=tol(real_t), atol=tol(real_t))
    # the CFL reduction leaves sum_i |u_i| per cell in the scalar buffer, as the reference
    dt_cfl = sim.compute_stable_timestep()
    mag = np.sum(np.fabs(u), axis=0)
    np.testing.assert_array_equal(to_host(sim.buffer_scalar_field), (np.fabs(u[0]) + np.fabs(u[1])) + np.fabs(u[2]))
    tol_t = 10 * np.finfo(real_t).eps
    want = min(0.1 * dx / (real_t(mag.max()) + tol_t), 0.9 * dx**2 / 6 / nu + tol_t)
    assert dt_cfl == want


@pytest.mark.parametrize("grid_size_x", [4, 8])
@pytest.mark.parametrize("precision", ["single", "double"])
def test_passive_transport_stable_timestep(dev, grid_size_x, precision):
    """test_passive_transport_flow_simulators.py:91-124: velocity 2 everywhere, cfl 0.2."""
    real_t = PREC[precision]
    nu, cfl, dt_prefac = 1e-2, 0.2, 0.5
    dx = real_t(1.0 / grid_size_x)
    sim = sps().PassiveTransportFlowSimulator(kinematic_viscosity=nu, cfl=cfl, grid_dim=3,
                                             grid_size=(grid_size_x,) * 3, x_range=1.0, real_t=real_t)
    sim.velocity_field.fill_(2.0)
    tol = 10 * np.finfo(real_t).eps
    ref = dt_prefac * min(cfl * dx / (3 * 2.0 + tol), 0.9 * dx**2 / (2 * 3) / nu)
    assert sim.compute_stable_timestep(dt_prefac=dt_prefac) == ref
