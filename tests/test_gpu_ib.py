"""GPU parity of the Eulerian-Lagrangian kernels and the virtual-boundary forcing against
the oracle. Spreading is checked bit-exact against the oracle's serial loop when both use
the same weights (the device reproduces the reference's per-cell summation order)."""

import numpy as np
import pytest

from helpers import PREC, rel_l2, to_dev, to_host, tol
from oracle import ib

pytestmark = pytest.mark.gpu


def comm_cls():
    from sopht_amd.numeric.immersed_boundary_ops import EulerianLagrangianGridCommunicator3D

    return EulerianLagrangianGridCommunicator3D


def vbf_cls():
    from sopht_amd.numeric.immersed_boundary_ops import VirtualBoundaryForcing

    return VirtualBoundaryForcing


def random_nodes(rng, m, lo=0.3, hi=0.7):
    return lo + (hi - lo) * rng.random((3, m))


@pytest.fixture(params=["single", "double"])
def real_t(request):
    return PREC[request.param]


@pytest.mark.parametrize("kind", ["cosine", "peskin"])
def test_support_and_weights(dev, real_t, kind, rng):
    n, m = 32, 257
    dx = real_t(1.0 / n)
    shift = real_t(dx / 2)
    X = random_nodes(rng, m)
    # include nodes exactly on grid points (floor edge cases)
    X[:, :3] = (np.arange(10, 13) * dx + shift)[None, :]
    idx_ref, d_ref = ib.local_eulerian_grid_support(X, dx, shift, real_t)
    f = ib.cosine_weights if kind == "cosine" else ib.peskin_weights
    w_ref, s_ref = f(d_ref, dx, real_t)
    c = comm_cls()(dx, shift, m, 2, real_t, n_components=3, interp_kernel_type=kind)
    import torch

    idx = torch.zeros((3, m), dtype=torch.int64, device=dev)
    sup = torch.zeros((3, 4, 4, 4, m), dtype=torch.float32 if real_t == np.float32 else torch.float64, device=dev)
    w = torch.zeros((4, 4, 4, m), dtype=sup.dtype, device=dev)
    c.local_eulerian_grid_support_of_lagrangian_grid_kernel(sup, idx, to_dev(X, dev))
    np.testing.assert_array_equal(to_host(idx), idx_ref)
    np.testing.assert_allclose(to_host(sup), d_ref, atol=tol(real_t), rtol=0)
    c.interpolation_weights_kernel(w, sup)
    np.testing.assert_allclose(to_host(sup), s_ref, atol=tol(real_t), rtol=0)
    np.testing.assert_allclose(to_host(w) * dx**3, w_ref * dx**3, atol=tol(real_t), rtol=0)


def test_interp_arange_weights_closed_form(dev, real_t):
    n, m = 16, 3
    dx = real_t(1.0 / n)
    idx = np.stack([np.arange(n // 3, n // 3 + m)] * 3).astype(np.int64)
    w = np.tile(np.arange(64, dtype=real_t).reshape(4, 4, 4, 1), (1, 1, 1, m))
    c = comm_cls()(dx, real_t(dx / 2), m, 2, real_t, n_components=3)
    u = to_dev(np.ones((3, n, n, n), dtype=real_t), dev)
    lag = to_dev(np.zeros((3, m), dtype=real_t), dev)
    c.eulerian_to_lagrangian_grid_interpolation_kernel(lag, u, to_dev(w, dev), to_dev(idx, dev))
    np.testing.assert_allclose(to_host(lag), 64 * 63 / 2 * dx**3, atol=tol(real_t))
    c1 = comm_cls()(dx, real_t(dx / 2), m, 2, real_t, n_components=1)
    lag1 = to_dev(np.zeros((m,), dtype=real_t), dev)
    c1.eulerian_to_lagrangian_grid_interpolation_kernel(lag1, u[0], to_dev(w, dev), to_dev(idx, dev))
    np.testing.assert_allclose(to_host(lag1), 64 * 63 / 2 * dx**3, atol=tol(real_t))


@pytest.mark.parametrize("m,cluster", [(3, False), (700, False), (500, True), (1500, True), (3000, False), (6000, False),
                                       (9000, False)])
def test_spread_bit_exact_vs_serial_order(dev, real_t, m, cluster, rng):
    n = 40
    dx = real_t(1.0 / n)
    shift = real_t(dx / 2)
    X = random_nodes(rng, m, 0.2, 0.8)
    if cluster:  # many nodes sharing a few cells (the worst case for an ordered gather)
        X = 0.5 + 0.3 * dx * rng.random((3, m))
    idx, d = ib.local_eulerian_grid_support(X, dx, shift, real_t)
    w, _ = ib.cosine_weights(d, dx, real_t)
    Fl = rng.standard_normal((3, m)).astype(real_t)
    F0 = rng.standard_normal((3, n, n, n)).astype(real_t)
    ref = F0.copy()
    ib.spread(ref, Fl, w, idx)
    c = comm_cls()(dx, shift, m, 2, real_t, n_components=3)
    F = to_dev(F0, dev)
    c.lagrangian_to_eulerian_grid_interpolation_kernel(F, to_dev(Fl, dev), to_dev(w, dev), to_dev(idx, dev))
    c.check_support_in_domain()
    np.testing.assert_array_equal(to_host(F), ref)
    # determinism: a second run from the same start gives identical bits
    F2 = to_dev(F0, dev)
    c.lagrangian_to_eulerian_grid_interpolation_kernel(F2, to_dev(Fl, dev), to_dev(w, dev), to_dev(idx, dev))
    np.testing.assert_array_equal(to_host(F2), to_host(F))


def test_spread_plan_reuse_and_rebuild(dev, real_t, rng):
    """One communicator (one gather plan) across calls: the same supports with new forces and
    weights (plan replayed), moved nodes (plan rebuilt), and a different grid whose packed
    keys would coincide (plan rebuilt) all stay bit-exact against the serial loop."""
    n, m = 32, 400
    dx = real_t(1.0 / n)
    shift = real_t(dx / 2)
    c = comm_cls()(dx, shift, m, 2, real_t, n_components=3)

    def check(X, shape, seed):
        r = np.random.default_rng(seed)
        idx, d = ib.local_eulerian_grid_support(X, dx, shift, real_t)
        w, _ = ib.cosine_weights(d, dx, real_t)
        w = (w * (1 + 0.1 * r.random(w.shape))).astype(real_t)  # weights change between calls too
        Fl = r.standard_normal((3, m)).astype(real_t)
        F0 = r.standard_normal((3,) + shape).astype(real_t)
        ref = F0.copy()
        ib.spread(ref, Fl, w, idx)
        F = to_dev(F0, dev)
        c.lagrangian_to_eulerian_grid_interpolation_kernel(F, to_dev(Fl, dev), to_dev(w, dev), to_dev(idx, dev))
        c.check_support_in_domain()
        np.testing.assert_array_equal(to_host(F), ref)

    X1 = random_nodes(rng, m, 0.25, 0.75)
    check(X1, (n, n, n), 1)
    check(X1, (n, n, n), 2)  # replay
    X2 = X1 + 1.7 * dx * (rng.random((3, m)) - 0.5)
    check(X2, (n, n, n), 3)  # rebuild
    check(X2, (n, n, n), 4)  # replay
    # same node coordinates on a grid with more z planes: every packed key agrees with the
    # previous call's, the cells do not
    check(X2, (n + 8, n, n), 5)
    check(X2, (n, n, n), 6)


def test_spread_declared_unchanged_support(dev, real_t, rng):
    """sopht_spread3d_reuse_plan (a body at rest): a declared call replays the plan without the
    sort and plan kernels and stays bit-exact; a declaration whose support changed is caught on
    the device (nothing is spread) and reported by the next check as ValueError; the call after
    it rebuilds the plan."""
    n, m = 32, 300
    dx = real_t(1.0 / n)
    shift = real_t(dx / 2)
    c = comm_cls()(dx, shift, m, 2, real_t, n_components=3)

    def call(X, seed, declare):
        r = np.random.default_rng(seed)
        idx, d = ib.local_eulerian_grid_support(X, dx, shift, real_t)
        w, _ = ib.cosine_weights(d, dx, real_t)
        Fl = r.standard_normal((3, m)).astype(real_t)
        F0 = r.standard_normal((3, n, n, n)).astype(real_t)
        ref = F0.copy()
        ib.spread(ref, Fl, w, idx)
        F = to_dev(F0, dev)
        if declare:
            c.declare_support_unchanged()
        c.lagrangian_to_eulerian_grid_interpolation_kernel(F, to_dev(Fl, dev), to_dev(w, dev), to_dev(idx, dev))
        return to_host(F), ref, F0

    X1 = random_nodes(rng, m, 0.25, 0.75)
    call(X1, 1, False)
    c.check_support_in_domain()
    got, ref, _ = call(X1, 2, True)  # same supports, declared
    c.check_support_in_domain()
    np.testing.assert_array_equal(got, ref)
    X2 = X1 + 2.5 * dx  # moved: the declaration is wrong
    got, _, F0 = call(X2, 3, True)
    np.testing.assert_array_equal(got, F0)  # nothing spread
    with pytest.raises(ValueError, match="reuse_plan"):
        c.check_support_in_domain()
    c.check_support_in_domain()  # cleared
    got, ref, _ = call(X2, 4, False)  # the plan is rebuilt
    c.check_support_in_domain()
    np.testing.assert_array_equal(got, ref)


def test_spread_scalar_and_out_of_domain(dev, rng):
    n, m = 16, 5
    real_t = np.float64
    dx = 1.0 / n
    X = random_nodes(rng, m)
    idx, d = ib.local_eulerian_grid_support(X, dx, dx / 2, real_t)
    w, _ = ib.cosine_weights(d, dx, real_t)
    Fl = rng.standard_normal(m)
    ref = np.zeros((n, n, n))
    ib.spread(ref, Fl, w, idx)
    c = comm_cls()(dx, dx / 2, m, 2, real_t, n_components=1)
    F = to_dev(np.zeros((n, n, n)), dev)
    c.lagrangian_to_eulerian_grid_interpolation_kernel(F, to_dev(Fl, dev), to_dev(w, dev), to_dev(idx, dev))
    np.testing.assert_array_equal(to_host(F), ref)
    # a node whose support leaves the grid is dropped and reported
    idx_bad = idx.copy()
    idx_bad[0, 0] = 0
    c.lagrangian_to_eulerian_grid_interpolation_kernel(F, to_dev(Fl, dev), to_dev(w, dev), to_dev(idx_bad, dev))
    with pytest.raises(IndexError):
        c.check_support_in_domain()
    c.check_support_in_domain()  # flag cleared


@pytest.mark.parametrize("kind", ["cosine", "peskin"])
@pytest.mark.parametrize("reset", [True, False])
def test_virtual_boundary_forcing_matches_oracle(dev, real_t, kind, reset, rng):
    n, m = 32, 400
    dx = real_t(1.0 / n)
    k, c = -5e4 * 1e-3, -10 * 1e-3
    X = random_nodes(rng, m)
    V = 0.1 * rng.standard_normal((3, m))
    u = rng.standard_normal((3, n, n, n)).astype(real_t)
    F0 = rng.standard_normal((3, n, n, n)).astype(real_t)
    o = ib.VirtualBoundaryForcing(k, c, dx, m, real_t, interp_kernel_type=kind, enable_eul_grid_forcing_reset=reset)
    o.lag_grid_position_mismatch_field[...] = 1e-3 * rng.standard_normal((3, m))
    g = vbf_cls()(k, c, 3, dx, m, real_t, enable_eul_grid_forcing_reset=reset, interp_kernel_type=kind)
    g.lag_grid_position_mismatch_field.copy_(to_dev(o.lag_grid_position_mismatch_field, dev))
    Fo = F0.copy()
    o.compute_interaction_forcing(Fo, u, X, V)
    Fd = to_dev(F0, dev)
    g.compute_interaction_forcing(Fd, to_dev(u, dev), X, V)
    np.testing.assert_array_equal(to_host(g.nearest_eul_grid_index_to_lag_grid), o.nearest_eul_grid_index_to_lag_grid)
    t = tol(real_t)
    np.testing.assert_allclose(to_host(g.interp_weights) * dx**3, o.interp_weights * dx**3, atol=t)
    np.testing.assert_allclose(to_host(g.lag_grid_flow_velocity_field), o.lag_grid_flow_velocity_field, atol=t)
    np.testing.assert_allclose(to_host(g.lag_grid_velocity_mismatch_field), o.lag_grid_velocity_mismatch_field, atol=t)
    np.testing.assert_allclose(to_host(g.lag_grid_forcing_field), o.lag_grid_forcing_field, atol=t)
    bound = 1e-12 if real_t == np.float64 else 1e-5
    assert rel_l2(to_host(Fd), Fo) < bound
    g.time_step(real_t(0.01))
    o.time_step(real_t(0.01))
    np.testing.assert_allclose(to_host(g.lag_grid_position_mismatch_field), o.lag_grid_position_mismatch_field, atol=t)
