"""GPU parity of the Neumann fast-diagonalisation Poisson solve (rocBLAS GEMMs) and of the
flow step with ``poisson_solver_type="fast_diagonalisation"``.

* The reference's closed-form tests (``test_unbounded_poisson_solver_3d.py:160-232``),
  scalar and vector, single and double, to 1e-4.
* Against the oracle (``numpy.linalg.eig`` + ``tensordot``) on random right-hand sides and
  anisotropic grids: relative L2 <= 1e-10 in FP64; FP32 against the FP64 oracle <= 2e-5.
* At 128^3 (FP64), beyond the oracle's sizes: the discrete identity
  ``-L_h psi = rhs - mean(rhs)``, ``mean(psi) = 0`` with an independent 7-point Neumann
  Laplacian.
* The eigenbasis the library builds against ``numpy.linalg.eig`` up to column signs;
  strided and aliased operands; the per-component path of the C entry point.
* One flow step of the simulator with the fast-diagonalisation solver against the oracle.
"""

import ctypes as C

import numpy as np
import pytest

from helpers import PREC, rel_l2, rel_linf, to_dev, to_host
from oracle import fastdiag as FD
from oracle import flow as OF

pytestmark = pytest.mark.gpu


def solver_cls():
    from sopht_amd.numeric.eulerian_grid_ops import FastDiagPoissonSolver3D

    return FastDiagPoissonSolver3D


def neumann_laplacian(f, dx):
    p = np.pad(f, 1, mode="edge")
    c = p[1:-1, 1:-1, 1:-1]
    lap = (p[2:, 1:-1, 1:-1] + p[:-2, 1:-1, 1:-1] + p[1:-1, 2:, 1:-1] + p[1:-1, :-2, 1:-1]
           + p[1:-1, 1:-1, 2:] + p[1:-1, 1:-1, :-2] - 6 * c)
    return lap / dx**2


def cos_rhs(n, real_t):
    dx = real_t(1.0 / n)
    x = np.linspace(dx / 2, 1 - dx / 2, n).astype(real_t)
    z, y, xx = np.meshgrid(x, x, x, indexing="ij")
    k = real_t(4)
    f = (np.cos(k * np.pi * xx) * np.cos(k * np.pi * y) * np.cos(k * np.pi * z)).astype(real_t)
    return f, dx, f / (3 * (k**2) * np.pi**2)


@pytest.mark.parametrize("precision", ["single", "double"])
def test_closed_form_scalar(dev, precision):
    """test_unbounded_poisson_solve_neumann_fast_diag_3d (:160-192)."""
    real_t = PREC[precision]
    rhs, dx, ref = cos_rhs(16, real_t)
    s = solver_cls()(16, 16, 16, dx=dx, real_t=real_t, bc_type="homogenous_neumann_along_xyz")
    out = to_dev(np.zeros_like(rhs), dev)
    s.solve(solution_field=out, rhs_field=to_dev(rhs, dev))
    assert np.amax(np.fabs(to_host(out) - ref)) <= real_t(1e-4)


@pytest.mark.parametrize("precision", ["single", "double"])
def test_closed_form_vector(dev, precision):
    """test_unbounded_poisson_solve_vector_field_neumann_fast_diag_3d (:195-232)."""
    real_t = PREC[precision]
    rhs, dx, ref = cos_rhs(16, real_t)
    rhs3 = np.stack([rhs] * 3)
    s = solver_cls()(16, 16, 16, dx=dx, real_t=real_t)
    out = to_dev(np.zeros_like(rhs3), dev)
    s.vector_field_solve(solution_vector_field=out, rhs_vector_field=to_dev(rhs3, dev))
    assert np.amax(np.fabs(to_host(out) - np.stack([ref] * 3))) <= real_t(1e-4)


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("shape", [(16, 16, 16), (8, 12, 20), (20, 6, 9), (2, 3, 5), (32, 24, 40)])
def test_vector_solve_matches_oracle(dev, precision, shape, rng):
    real_t = PREC[precision]
    dx = 0.05
    rhs = rng.standard_normal((3, *shape)).astype(real_t)
    ref = np.zeros(rhs.shape)
    FD.FastDiagPoissonSolver3D(*shape, dx).vector_field_solve(ref, rhs.astype(np.float64))
    s = solver_cls()(*shape, dx=dx, real_t=real_t)
    out = to_dev(np.zeros_like(rhs), dev)
    s.vector_field_solve(out, to_dev(rhs, dev))
    got = to_host(out)
    bound = 1e-10 if real_t == np.float64 else 2e-5
    assert rel_l2(got, ref) <= bound, rel_l2(got, ref)
    assert rel_linf(got, ref) <= 5 * bound, rel_linf(got, ref)


@pytest.mark.parametrize("precision", ["single", "double"])
def test_scalar_solve_matches_oracle_same_precision(dev, precision, rng):
    """The oracle in the solver's own precision (eig in FP32 for single, as the reference)."""
    real_t = PREC[precision]
    shape = (12, 16, 10)
    dx = real_t(0.1)
    rhs = rng.standard_normal(shape).astype(real_t)
    ref = np.zeros_like(rhs)
    FD.FastDiagPoissonSolver3D(*shape, dx, real_t).solve(ref, rhs)
    s = solver_cls()(*shape, dx=dx, real_t=real_t)
    out = to_dev(np.zeros_like(rhs), dev)
    s.solve(out, to_dev(rhs, dev))
    bound = 1e-10 if real_t == np.float64 else 2e-5
    assert rel_l2(to_host(out), ref) <= bound, rel_l2(to_host(out), ref)


def test_discrete_identity_128(dev, rng):
    import torch

    n = 128
    dx = 1.0 / n
    rhs = torch.randn((3, n, n, n), dtype=torch.float64, device=dev)
    out = torch.empty_like(rhs)
    solver_cls()(n, n, n, dx=dx, real_t=np.float64).vector_field_solve(out, rhs)
    for c in range(3):
        r = to_host(rhs[c])
        psi = to_host(out[c])
        res = -neumann_laplacian(psi, dx) - (r - r.mean())
        assert np.abs(res).max() <= 1e-9 * np.abs(r).max(), np.abs(res).max()
        assert abs(psi.mean()) <= 1e-12 * np.abs(psi).max()


def test_eigenbasis_matches_eig(dev):
    shape = (7, 12, 16)
    dx = 0.2
    s = solver_cls()(*shape, dx=dx, real_t=np.float64)
    for name, n in zip("zyx", shape):
        vals, vecs, _ = FD.sorted_eig(FD.neumann_poisson_matrix(n, dx, np.float64))
        got_vals = getattr(s, f"eig_vals_{name}")
        got = getattr(s, f"eig_vecs_{name}")
        np.testing.assert_allclose(got_vals, vals, rtol=0, atol=1e-12 / dx**2)
        sign = np.sign(np.sum(got * vecs, axis=0))
        np.testing.assert_allclose(vecs * sign, got, rtol=0, atol=1e-12)
        np.testing.assert_allclose(getattr(s, f"inv_of_eig_vecs_{name}") @ got, np.eye(n), atol=1e-13)
    inv = to_host(s.inv_eig_val_matrix)
    assert inv[-1, -1, -1] == 0 and np.all(np.isfinite(inv))


def test_strided_and_aliased(dev, rng):
    import torch

    shape = (10, 12, 14)
    dx = 0.1
    s = solver_cls()(*shape, dx=dx, real_t=np.float64)
    rhs = rng.standard_normal(shape)
    ref = np.zeros(shape)
    FD.FastDiagPoissonSolver3D(*shape, dx).solve(ref, rhs)
    # non-contiguous operands (views of a larger tensor)
    big = torch.zeros((2, 12, 14, 18), dtype=torch.float64, device=dev)
    big[1, 1:11, 1:13, 2:16] = to_dev(rhs, dev)
    rv = big[1, 1:11, 1:13, 2:16]
    sbig = torch.zeros_like(big)
    sv = sbig[0, 2:12, 0:12, 1:15]
    s.solve(sv, rv)
    assert rel_l2(to_host(sv), ref) <= 1e-10
    assert float(sbig.abs().sum()) == pytest.approx(float(sv.abs().sum()))
    np.testing.assert_array_equal(to_host(rv), rhs)
    # solution aliasing rhs
    t = to_dev(rhs, dev)
    s.solve(t, t)
    assert rel_l2(to_host(t), ref) <= 1e-10


def test_per_component_c_path(dev, rng):
    """Components that are separate allocations take the per-component sequence."""
    import torch

    from sopht_amd import _lib as L

    shape = (8, 10, 12)
    dx = 0.1
    s = solver_cls()(*shape, dx=dx, real_t=np.float64)
    rhs = rng.standard_normal((3, *shape))
    ref = np.zeros_like(rhs)
    FD.FastDiagPoissonSolver3D(*shape, dx).vector_field_solve(ref, rhs)
    ins = [to_dev(rhs[c], dev) for c in (2, 0, 1)]
    outs = [torch.zeros(shape, dtype=torch.float64, device=dev) for _ in range(3)]
    order = (1, 2, 0)
    sol = L._p3(*(outs[i].data_ptr() for i in order))
    rp = L._p3(*(ins[i].data_ptr() for i in order))
    L.call("sopht_fastdiag3d_vector_solve", s.plan, sol, L.strides3(outs[0]), rp,
           L.strides3(ins[0]), L.stream_ptr())
    for c, i in enumerate(order):
        assert rel_l2(to_host(outs[i]), ref[c]) <= 1e-10


@pytest.mark.parametrize("precision", ["single", "double"])
def test_flow_step_with_fast_diagonalisation(dev, precision, rng):
    """poisson_solver_type="fast_diagonalisation" (navier_stokes_flow_simulators.py:349-358):
    one forced step with free stream against the oracle."""
    from sopht_amd import simulator as sps

    real_t = PREC[precision]
    shape = (16, 12, 20)
    kw = dict(kinematic_viscosity=1e-2, time=0.0, with_forcing=True, with_free_stream_flow=True)
    o = OF.UnboundedNavierStokesFlowSimulator3D(shape, 1.0, real_t=np.float64,
                                               poisson_solver_type="fast_diagonalisation", **kw)
    g = sps.UnboundedNavierStokesFlowSimulator3D(grid_size=shape, x_range=1.0, real_t=real_t,
                                                poisson_solver_type="fast_diagonalisation", **kw)
    w0, u0, F0 = (rng.random((3,) + shape).astype(real_t) for _ in range(3))
    o.vorticity_field[...] = w0
    o.velocity_field[...] = u0
    o.eul_grid_forcing_field[...] = F0
    g.vorticity_field.copy_(to_dev(w0, dev))
    g.velocity_field.copy_(to_dev(u0, dev))
    g.eul_grid_forcing_field.copy_(to_dev(F0, dev))
    U = np.array([1.0, 0.5, -0.25])
    for _ in range(2):
        o.time_step(dt=1e-3, free_stream_velocity=U)
        g.time_step(dt=1e-3, free_stream_velocity=U)
    bound = 1e-10 if real_t == np.float64 else 2e-5
    for name in ("vorticity_field", "velocity_field", "stream_func_field"):
        a, b = to_host(getattr(g, name)), getattr(o, name)
        assert rel_l2(a, b) <= bound, (name, rel_l2(a, b))


def test_fast_diagonalisation_is_rejected_by_the_adjoint(dev):
    from sopht_amd import _lib as L
    from sopht_amd import simulator as sps

    g = sps.UnboundedNavierStokesFlowSimulator3D(grid_size=(8, 8, 8), x_range=1.0, real_t=np.float64,
                                                kinematic_viscosity=1e-2, with_forcing=True,
                                                poisson_solver_type="fast_diagonalisation")
    import torch

    z = torch.zeros((3, 8, 8, 8), dtype=torch.float64, device=dev)
    with pytest.raises(ValueError):
        L.call("sopht_ns3d_adjoint_step", g._plan, L.ptr(z), L.ptr(z), 1e-3, L.ptr(z), L.ptr(z),
               L.ptr(z), L.stream_ptr())
