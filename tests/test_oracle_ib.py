"""Pins the oracle's Eulerian-Lagrangian transfer and virtual-boundary forcing (CPU only).

Closed forms follow the reference tests
``tests/test_numeric/test_immersed_boundary_ops/test_eulerian_lagrangian_grid_communicator_3d.py``
(support at grid points ``:7-117``, interpolation with weights ``arange(64)`` ``:120-231``,
spreading mass conservation ``:234-369``, kernel value at offset (0, dx, dx) ``:372-421``) and
``test_virtual_boundary_forcing.py`` (mock solution ``:7-100``, ``time_step`` ``:383-420``),
plus partition of unity and the transpose identity between interpolation and spreading.
"""

import numpy as np
import pytest

from oracle import ib

PRECISIONS = [np.float32, np.float64]


def tol(real_t):
    return real_t(1e3) * np.finfo(real_t).eps


def grid_nodes(n=16, m=3):
    idx = np.empty((3, m), dtype=np.int64)
    for a in range(3):
        idx[a] = np.arange(n // 3, n // 3 + m)
    return idx


@pytest.mark.parametrize("real_t", PRECISIONS)
def test_support_at_grid_points(real_t):
    n = 16
    dx = real_t(1.0 / n)
    shift = real_t(0.5 * dx)
    idx_ref = grid_nodes(n)
    X = idx_ref * dx + shift
    idx, d = ib.local_eulerian_grid_support(X, dx, shift, real_t)
    np.testing.assert_array_equal(idx, idx_ref)
    # support = coordinates of the 4^3 block minus the node position
    c = (np.arange(n) * dx + shift).astype(real_t)
    for i in range(3):
        blk = [slice(idx_ref[a, i] - 1, idx_ref[a, i] + 3) for a in range(3)]
        np.testing.assert_allclose(d[0, :, :, :, i], np.broadcast_to(c[blk[0]][None, None, :], (4, 4, 4)) - X[0, i], atol=tol(real_t))
        np.testing.assert_allclose(d[1, :, :, :, i], np.broadcast_to(c[blk[1]][None, :, None], (4, 4, 4)) - X[1, i], atol=tol(real_t))
        np.testing.assert_allclose(d[2, :, :, :, i], np.broadcast_to(c[blk[2]][:, None, None], (4, 4, 4)) - X[2, i], atol=tol(real_t))


@pytest.mark.parametrize("real_t", PRECISIONS)
def test_interpolation_arange_weights(real_t):
    n = 16
    dx = real_t(1.0 / n)
    idx = grid_nodes(n)
    w = np.tile(np.arange(64, dtype=real_t).reshape(4, 4, 4, 1), (1, 1, 1, 3))
    u = np.ones((3, n, n, n), dtype=real_t)
    out = ib.interpolate(u, w, idx, dx)
    np.testing.assert_allclose(out, 64 * 63 / 2 * dx**3, atol=tol(real_t))


@pytest.mark.parametrize("real_t", PRECISIONS)
def test_spread_unit_weights_mass(real_t):
    n = 16
    idx = grid_nodes(n)
    w = np.ones((4, 4, 4, 3), dtype=real_t)
    F = np.zeros((3, n, n, n), dtype=real_t)
    ib.spread(F, 2 * np.ones((3, 3), dtype=real_t), w, idx)
    np.testing.assert_allclose(F.sum(axis=(1, 2, 3)), 3 * 2 * 64, atol=tol(real_t))


@pytest.mark.parametrize("real_t", PRECISIONS)
@pytest.mark.parametrize("kind", ["cosine", "peskin"])
def test_kernel_value_at_offset(real_t, kind):
    dx = real_t(1.0 / 16)
    d = np.zeros((3, 4, 4, 4, 3), dtype=real_t)
    d[1:] = dx
    f = ib.cosine_weights if kind == "cosine" else ib.peskin_weights
    w, _ = f(d, dx, real_t)
    np.testing.assert_allclose(w, real_t(2 * (0.25 / dx) ** 3), atol=tol(real_t))


@pytest.mark.parametrize("kind", ["cosine", "peskin"])
def test_partition_of_unity_and_transpose(kind):
    rng = np.random.default_rng(11)
    n, m = 20, 37
    dx = 1.0 / n
    shift = dx / 2
    X = 0.3 + 0.4 * rng.random((3, m))
    idx, d = ib.local_eulerian_grid_support(X, dx, shift, np.float64)
    f = ib.cosine_weights if kind == "cosine" else ib.peskin_weights
    w, _ = f(d, dx, np.float64)
    # sum_block w dx^3 = 1 for any node position
    np.testing.assert_allclose(w.sum(axis=(0, 1, 2)) * dx**3, 1.0, atol=1e-13)
    # interpolating a constant field returns the constant; a linear field is reproduced
    # exactly by the cosine kernel's first moment (symmetric kernel)
    u = np.full((3, n, n, n), 2.5)
    np.testing.assert_allclose(ib.interpolate(u, w, idx, dx), 2.5, atol=1e-12)
    # <interp(u), F> = dx^3 <u, spread(F)>
    u = rng.standard_normal((3, n, n, n))
    Fl = rng.standard_normal((3, m))
    lhs = np.sum(ib.interpolate(u, w, idx, dx) * Fl)
    Fe = np.zeros_like(u)
    ib.spread(Fe, Fl, w, idx)
    rhs = dx**3 * np.sum(u * Fe)
    assert abs(lhs - rhs) < 1e-12 * max(1.0, abs(lhs))


@pytest.mark.parametrize("real_t", PRECISIONS)
@pytest.mark.parametrize("reset", [True, False])
def test_virtual_boundary_mock_solution(real_t, reset):
    n = 16
    k, c = real_t(1e3), real_t(1e1)
    dx = real_t(1.0 / n)
    vb = ib.VirtualBoundaryForcing(k, c, dx, 3, real_t, enable_eul_grid_forcing_reset=reset)
    idx = grid_nodes(n)
    X = idx * dx + real_t(dx / 2)
    u = real_t(5.0) * np.ones((3, n, n, n), dtype=real_t)
    V = real_t(2.0) * np.ones((3, 3), dtype=real_t)
    F = np.full((3, n, n, n), 0.0 if reset else 0.0, dtype=real_t)
    vb.compute_interaction_forcing(F, u, X, V)
    np.testing.assert_array_equal(vb.nearest_eul_grid_index_to_lag_grid, idx)
    np.testing.assert_allclose(vb.lag_grid_flow_velocity_field, 5.0, atol=tol(real_t))
    np.testing.assert_allclose(vb.lag_grid_velocity_mismatch_field, 3.0, atol=tol(real_t))
    np.testing.assert_allclose(vb.lag_grid_forcing_field, 30.0, atol=10 * tol(real_t))
    # force conservation: sum F_eul dx^3 == sum F_lag
    np.testing.assert_allclose(F.sum() * dx**3, vb.lag_grid_forcing_field.sum(), rtol=1e-5 if real_t == np.float32 else 1e-12)
    assert np.sum(F > 0) <= 3 * 3 * 64
    # time_step: dX += dt dV
    vb.time_step(real_t(0.1))
    np.testing.assert_allclose(vb.lag_grid_position_mismatch_field, 0.3, atol=tol(real_t))
    assert vb.time == pytest.approx(0.1)
