"""GPU parity of the reverse-mode adjoint (``sopht_ns3d_adjoint_step``, ``sopht_vbf_adjoint_3d``,
the checkpointed rollout behind ``torch.autograd.Function``) against the oracle's adjoint,
which tests/test_oracle_adjoint.py pins by transpose identities and finite differences.

* out-of-place trajectory step == in-place step, bit for bit; forcing left intact;
* one flow-step adjoint and one virtual-boundary adjoint vs the oracle (FP64, 1e-12);
* the whole rollout gradient vs the oracle's (FP64, 1e-9), across checkpoint segments;
* the gradient of the time-averaged body x-force of a heaving ellipsoid (64^3 FP64, 200
  steps, 2,000 nodes, 16 Fourier coefficients) vs central finite differences with h = 1e-4
  on the same frozen dt sequence: relative L2 error <= 1e-4;
* FP32 gradient vs FP64 gradient (reported, bounded).
"""

import ctypes as C

import numpy as np
import pytest

from helpers import PREC, rel_l2, to_dev, to_host
from oracle import adjoint as OA
from oracle import flow as OF
from oracle import ib as OIB
from sopht_amd.utils import synthetic

pytestmark = pytest.mark.gpu


def sps():
    from sopht_amd import simulator

    return simulator


def L():
    from sopht_amd import _lib

    return _lib


def gpu_sim(shape, real_t, nu=1e-2, rho=1.0):
    return sps().UnboundedNavierStokesFlowSimulator3D(
        grid_size=shape, x_range=1.0, kinematic_viscosity=nu, real_t=real_t, with_forcing=True,
        with_free_stream_flow=True, flow_density=rho)


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("shape", [(12, 8, 16), (8, 12, 6)])
def test_trajectory_step_matches_in_place_step(dev, precision, shape, rng):
    import torch

    real_t = PREC[precision]
    g = gpu_sim(shape, real_t)
    w0, u0, F0 = (to_dev(rng.random((3,) + shape).astype(real_t), dev) for _ in range(3))
    W1 = torch.empty_like(w0)
    U1 = torch.empty_like(w0)
    F = F0.clone()
    U = np.array([1.0, -0.5, 0.25])
    w1 = torch.empty_like(w0)
    g.trajectory_step(w0, u0, F, W1, U1, 1e-3, U, vorticity1_out=w1)
    assert torch.equal(F, F0)
    # omega1 = omega + dt/(2 dx rho) curl_h F on the interior (the oracle's forcing update)
    o = OF.UnboundedNavierStokesFlowSimulator3D(shape, 1.0, 1e-2, real_t=np.float64, with_forcing=True)
    ref = to_host(w0).astype(np.float64)
    from oracle import stencils as OS

    OS.update_vorticity_from_velocity_forcing(ref, to_host(F0).astype(np.float64),
                                              np.float64(1e-3 / (2 * float(g.dx))))
    assert rel_l2(to_host(w1), ref) <= (1e-15 if real_t == np.float64 else 1e-6)
    g.vorticity_field.copy_(w0)
    g.velocity_field.copy_(u0)
    g.eul_grid_forcing_field.copy_(F0)
    g.time_step(dt=1e-3, free_stream_velocity=U)
    assert torch.equal(W1, g.vorticity_field)
    assert torch.equal(U1, g.velocity_field)


@pytest.mark.parametrize("shape", [(12, 8, 16), (8, 12, 6), (16, 16, 16), (24, 16, 12)])
@pytest.mark.parametrize("rho", [1.0, 0.5])
@pytest.mark.parametrize("width", [2, 1, 3])
def test_flow_step_adjoint_matches_oracle(dev, shape, rho, width, rng):
    import torch

    real_t = np.float64
    nu, dt = 1e-2, 2e-3
    o = OF.UnboundedNavierStokesFlowSimulator3D(shape, 1.0, nu, real_t=real_t, with_forcing=True,
                                                with_free_stream_flow=True, flow_density=rho,
                                                penalty_zone_width=width)
    g = sps().UnboundedNavierStokesFlowSimulator3D(
        grid_size=shape, x_range=1.0, kinematic_viscosity=nu, real_t=real_t, with_forcing=True,
        with_free_stream_flow=True, flow_density=rho, penalty_zone_width=width)
    w, u, F, wb, ub = (rng.standard_normal((3,) + shape) for _ in range(5))
    ow, ou, oF = OA.flow_step_adjoint(o, dt, w, u, F, wb, ub)
    w1 = w.copy()  # omega1, the adjoint step's input
    from oracle import stencils as OS

    OS.update_vorticity_from_velocity_forcing(w1, F, OA.step_prefactors(o, dt)[0])
    d = [to_dev(a, dev) for a in (w1, u, wb, ub)]
    work = torch.empty((3,) + shape, dtype=torch.float64, device=dev)
    L().call("sopht_ns3d_adjoint_step", g._plan, L().ptr(d[0]), L().ptr(d[1]), dt, L().ptr(d[2]),
             L().ptr(d[3]), L().ptr(work), L().stream_ptr())
    assert rel_l2(to_host(d[2]), ow) <= 1e-12, rel_l2(to_host(d[2]), ow)
    assert rel_l2(to_host(d[3]), ou) <= 1e-12, rel_l2(to_host(d[3]), ou)


def test_vbf_adjoint_matches_oracle(dev, rng):
    import torch

    shape = (16, 24, 12)
    nz, ny, nx = shape
    dx = 1.0 / nx
    shift = dx / 2
    n = 37
    X = 0.3 + 0.4 * rng.random((3, n)) * np.array([[1.0], [0.8], [0.7]])
    u, wb, ub = (rng.standard_normal((3,) + shape) for _ in range(3))
    Fl, dXb, dVb = (rng.standard_normal((3, n)) for _ in range(3))
    fbar = rng.standard_normal(3)
    k, c, dt, pF = -40.0, -2.5, 1.5e-3, 0.37
    idx, d = OIB.local_eulerian_grid_support(X, dx, shift, np.float64)
    w, _ = OIB.cosine_weights(d, dx, np.float64)
    Fbar = OA.curl_transpose(wb, pF)
    ub_ref = ub.copy()
    Xb_ref, Vb_ref, dXb_ref, dVb_ref = OA.vbf_adjoint(Fbar, u, X, Fl, idx, w, fbar, dXb, dVb, k, c,
                                                       dt, dx, shift, ub_ref)
    t = lambda a: to_dev(np.ascontiguousarray(a, dtype=np.float64), dev)  # noqa: E731
    dX, dV, gub = t(dXb), t(dVb), t(ub)
    # keep every device input referenced until the kernels have run
    gwb, gu, gX, gFl, gfbar = t(wb), t(u), t(X), t(Fl), t(fbar)
    Xb = torch.empty((3, n), dtype=torch.float64, device=dev)
    Vb = torch.empty_like(Xb)
    wts = torch.empty((64, n), dtype=torch.float64, device=dev)
    gidx = torch.empty((3, n), dtype=torch.int64, device=dev)
    lag = torch.empty((3, n), dtype=torch.float64, device=dev)
    ws = C.c_void_p()
    L().call("sopht_spread3d_create", n, C.byref(ws))
    try:
        L().call("sopht_vbf_adjoint_3d", L().SOPHT_F64, L().ptr(gwb), pF, L().ptr(gu), nz, ny, nx,
                 L().ptr(gX), L().ptr(gFl), n, dx, shift, k, c, dt, L().ptr(gfbar), L().ptr(dX),
                 L().ptr(dV), L().ptr(Xb), L().ptr(Vb), L().ptr(wts), L().ptr(gidx), L().ptr(lag), ws,
                 L().ptr(gub), L().stream_ptr())
        L().call("sopht_spread3d_check", ws, L().stream_ptr())
    finally:
        L().lib().sopht_spread3d_destroy(ws)
    for a, b in ((Xb, Xb_ref), (Vb, Vb_ref), (dX, dXb_ref), (dV, dVb_ref), (gub, ub_ref)):
        assert rel_l2(to_host(a), b) <= 1e-12, rel_l2(to_host(a), b)


def build_rollout(shape, real_t, n_nodes, semi_axes, nu, period, num_steps, checkpoint_every,
                  omega0=None, k_coef=-5e4, c_coef=-10.0):
    """GPU rollout of a heaving Fibonacci ellipsoid at the domain centre; the stiffness and
    damping are rescaled by the maximum Lagrangian spacing squared as the interaction does
    (immersed_body_flow_interaction.py:83-86)."""
    S = sps()
    from sopht_amd.numeric.immersed_boundary_ops import VirtualBoundaryForcing

    flow = gpu_sim(shape, real_t, nu=nu)
    U = np.array([1.0, 0.0, 0.0])
    if omega0 is not None:
        flow.vorticity_field.copy_(to_dev(omega0.astype(real_t), flow.vorticity_field.device))
    flow.recover_velocity(U)
    X0 = synthetic.fibonacci_ellipsoid(n_nodes, semi_axes, (0.5, 0.5, 0.5))
    body = S.RigidBody((0.5, 0.5, 0.5))
    grid = S.FibonacciEllipsoidForcingGrid(3, body, n_nodes, semi_axes)
    ds = grid.get_maximum_lagrangian_grid_spacing()
    k, c = k_coef * ds**2, c_coef * ds**2
    inter = VirtualBoundaryForcing(k, c, 3, flow.dx, n_nodes, real_t, enable_eul_grid_forcing_reset=False)
    kin = S.HeaveKinematics(X0, period)
    ro = S.CoupledFlowRollout(flow, inter, kin, U, num_steps, checkpoint_every=checkpoint_every)
    return ro, X0, k, c, U


@pytest.mark.parametrize("step_workspaces", [True, False])
def test_rollout_gradient_matches_oracle(dev, step_workspaces, monkeypatch):
    """With one spreading workspace per recorded step (the backward replays the forward's
    plans) and, past the workspace budget, per step of a segment (the recompute rebuilds)."""
    import torch

    if not step_workspaces:
        monkeypatch.setattr(sps().CoupledFlowRollout, "STEP_WS_BUDGET", 0)
    shape = (16, 16, 16)
    nu, period, ns = 1e-2, 0.02, 6
    omega0 = synthetic.random_phase_vorticity(shape, 1.0, max_wavenumber=3, seed=3, max_vorticity=4.0)
    ro, X0, k, c, U = build_rollout(shape, np.float64, 40, (0.2, 0.12, 0.12), nu, period, ns, 4,
                                    omega0=omega0, k_coef=-60.0, c_coef=-3.0)
    assert (len(ro._step_ws) == ns) == step_workspaces
    theta = torch.tensor(np.random.default_rng(7).normal(0.0, 0.5, 16), device=dev, requires_grad=True)
    forces = sps().body_forces(theta, ro)
    J = sps().time_averaged_force(forces, ro.dts)
    J.backward()
    # oracle on the same dt sequence
    o = OA.Rollout(shape, 1.0, nu, omega0, U, OA.HeaveKinematics(X0, period), k=k, c=c)
    of, _ = o.forward(theta.detach().cpu().numpy(), ns, dts=ro.dts)
    assert rel_l2(to_host(forces), of) <= 1e-10, rel_l2(to_host(forces), of)
    fbar = np.zeros((ns, 3))
    fbar[:, 0] = OA.time_average_weights(ro.dts)
    og = o.gradient(theta.detach().cpu().numpy(), ns, ro.dts, fbar)
    err = rel_l2(to_host(theta.grad), og)
    print(f"rollout gradient vs oracle: rel L2 {err:.3e}")
    assert err <= 1e-9, err


def test_gradient_matches_finite_differences_64(dev):
    """64^3 FP64, 200 steps, ellipsoid (0.1, 0.05, 0.05) with 2,000 nodes, theta ~ N(0, 0.1)
    (seed 5), objective the time-averaged body x-force, central differences h = 1e-4 on the
    frozen dt sequence of the base run, checkpoints every 20 steps."""
    import time

    import torch

    shape = (64, 64, 64)
    ns = 200
    ro, *_ = build_rollout(shape, np.float64, 2000, (0.1, 0.05, 0.05), 1e-3, 0.25, ns, 20)
    theta = torch.tensor(np.random.default_rng(5).normal(0.0, 0.1, 16), device=dev, requires_grad=True)
    t0 = time.perf_counter()
    forces = sps().body_forces(theta, ro)
    J = sps().time_averaged_force(forces, ro.dts)
    J.backward()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    g = to_host(theta.grad)
    dts = ro.dts.copy()
    h = 1e-4

    def Jat(th):
        f = ro.run(torch.tensor(th, device=dev), dt_schedule=dts, record=False)
        return float(sps().time_averaged_force(f, dts))

    th0 = theta.detach().cpu().numpy()
    fd = np.array([(Jat(th0 + h * e) - Jat(th0 - h * e)) / (2 * h) for e in np.eye(16)])
    err = rel_l2(g, fd)
    print(f"64^3 FP64 adjoint vs FD: rel L2 {err:.3e}; fwd+adj {t1 - t0:.2f} s; |g| {np.linalg.norm(g):.3e}")
    assert np.linalg.norm(fd) > 0
    assert err <= 1e-4, (err, g, fd)


def test_fp32_gradient_close_to_fp64(dev):
    import torch

    shape = (32, 32, 32)
    ns = 40
    grads = {}
    dts = None
    for real_t in (np.float64, np.float32):
        ro, *_ = build_rollout(shape, real_t, 600, (0.15, 0.08, 0.08), 1e-3, 0.1, ns, 20)
        theta = torch.tensor(np.random.default_rng(11).normal(0.0, 0.1, 16), device=dev, requires_grad=True)
        forces = sps().body_forces(theta, ro, dts)
        if dts is None:
            dts = ro.dts.copy()
        sps().time_averaged_force(forces, dts).backward()
        grads[real_t] = to_host(theta.grad)
    err = rel_l2(grads[np.float32], grads[np.float64])
    print(f"FP32 vs FP64 gradient (32^3, 40 steps): rel L2 {err:.3e}")
    assert err <= 1e-2, err
