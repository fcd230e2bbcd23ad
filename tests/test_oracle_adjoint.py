"""Pins the oracle's discrete adjoint of the coupled flow rollout (CPU only).

The reference has no adjoint, so nothing in it fixes these functions; the mathematics does:
* each transpose satisfies ``<A x, y> = <x, A^T y>`` against the oracle's forward operators
  (which the other oracle tests pin to the reference) on random inputs;
* the cosine weight gradient matches central differences of the weights;
* the rollout gradient (dt frozen) matches central finite differences of the rollout.
A dropped term, a wrong sign or a transposed index in any transpose fails one of these.
"""

import numpy as np
import pytest

from oracle import adjoint as A
from oracle import flow as F
from oracle import ib, poisson, stencils
from sopht_amd.utils import synthetic


def dot(a, b):
    return float(np.sum(np.asarray(a, dtype=np.float64) * np.asarray(b, dtype=np.float64)))


def check_transpose(fwd, adj, x, y, rtol=1e-12):
    lhs = dot(fwd(x), y)
    rhs = dot(x, adj(y))
    assert abs(lhs - rhs) <= rtol * max(abs(lhs), abs(rhs), 1e-300), (lhs, rhs)


SHAPES = [(8, 9, 10), (12, 7, 6)]


@pytest.mark.parametrize("shape", SHAPES)
def test_curl_transpose(rng, shape):
    p = 0.37

    def fwd(f):
        out = rng.standard_normal(f.shape) * 0  # border reset to 0 (curl_3d.py:103-130)
        stencils.curl(out, f, p)
        return out

    check_transpose(fwd, lambda g: A.curl_transpose(g, p), rng.standard_normal((3, *shape)),
                    rng.standard_normal((3, *shape)))


@pytest.mark.parametrize("shape", SHAPES)
def test_diffusion_transpose(rng, shape):
    pD = 0.11

    def fwd(w):
        w = w.copy()
        stencils.diffusion_timestep_euler_forward(w, np.zeros(shape), pD)
        return w

    check_transpose(fwd, lambda g: A.diffusion_transpose(g, pD), rng.standard_normal((3, *shape)),
                    rng.standard_normal((3, *shape)))


@pytest.mark.parametrize("shape", [(8, 9, 10), (6, 6, 6)])
@pytest.mark.parametrize("width", [1, 2, 3])
def test_penalise_transpose(rng, shape, width):
    flow = F.UnboundedNavierStokesFlowSimulator3D(shape, 1.0, 1e-3, penalty_zone_width=width)
    mats = A.penalise_matrices(flow)
    x = rng.standard_normal((3, *shape))
    # the separable matrices reproduce the reference's in-place sequence
    ref = x.copy()
    flow.penalise(ref)
    np.testing.assert_allclose(A.penalise_apply(x, mats), ref, rtol=0, atol=1e-15)
    check_transpose(lambda f: A.penalise_apply(f, mats),
                    lambda g: A.penalise_apply(g, mats, transpose=True), x,
                    rng.standard_normal((3, *shape)))


def test_cross_transposes(rng):
    shape = (5, 6, 7)
    u, a, g = (rng.standard_normal((3, *shape)) for _ in range(3))
    # B = u x a: d/du <g, B> = a x g ; d/da <g, B> = g x u
    check_transpose(lambda du: A.cross(du, a), lambda gg: A.cross(a, gg), u, g)
    check_transpose(lambda da: A.cross(u, da), lambda gg: A.cross(gg, u), a, g)
    # the component formulas are numpy's cross product
    np.testing.assert_allclose(A.cross(u, a), np.cross(u, a, axis=0), rtol=1e-14, atol=1e-14)


def test_poisson_symmetric(rng):
    shape = (8, 10, 6)
    solver = poisson.UnboundedPoissonSolver3D(*shape, x_range=1.0)

    def G(f):
        out = np.zeros_like(f)
        solver.solve(out, f)
        return out

    check_transpose(G, G, rng.standard_normal(shape), rng.standard_normal(shape), rtol=1e-12)


def test_cosine_weight_gradient(rng):
    n = 7
    dx = 1.0 / 16
    shift = dx / 2
    X = 0.3 + 0.4 * rng.random((3, n))
    idx, _ = ib.local_eulerian_grid_support(X, dx, shift, np.float64)
    dw = A.cosine_weight_gradient(idx, X, dx, shift)
    h = 1e-7
    for a in range(3):
        e = np.zeros((3, 1))
        e[a] = h
        # keep the block fixed (idx) while differentiating: weights at fixed cells
        wp = _weights_fixed_idx(idx, X + e, dx, shift)
        wm = _weights_fixed_idx(idx, X - e, dx, shift)
        np.testing.assert_allclose(dw[a], (wp - wm) / (2 * h), rtol=1e-6, atol=1e-6 * np.abs(dw).max())


def _weights_fixed_idx(idx, X, dx, shift):
    n = X.shape[1]
    off = ib.support_offsets()
    oz, oy, ox = np.meshgrid(off, off, off, indexing="ij")
    offs = np.stack((ox, oy, oz)).reshape(3, 4, 4, 4, 1)
    d = (idx.reshape(3, 1, 1, 1, n) + offs) * dx + shift - X.reshape(3, 1, 1, 1, n)
    w, _ = ib.cosine_weights(d, dx, np.float64)
    return w


def test_weight_gradient_continuous_across_cells():
    # a node crossing a cell face changes the block; the cosine kernel keeps w and dw/dX
    # continuous there (the cell that leaves the block has w = dw = 0), which the
    # finite-difference check of the rollout relies on
    dx = 1.0 / 16
    shift = dx / 2
    x_face = 8 * dx + shift
    per_cell = []
    for side in (-1e-11, 1e-11):
        X = np.array([[x_face + side], [0.5], [0.5]])
        idx, d = ib.local_eulerian_grid_support(X, dx, shift, np.float64)
        w, _ = ib.cosine_weights(d, dx, np.float64)
        dw = A.cosine_weight_gradient(idx, X, dx, shift)
        cells = idx[0, 0] + ib.support_offsets()
        per_cell.append({int(c): (w[:, :, k, 0], dw[0][:, :, k, 0]) for k, c in enumerate(cells)})
    left, right = per_cell
    scale = max(np.abs(v[1]).max() for v in left.values())
    for c in set(left) | set(right):
        wl, dl = left.get(c, (0.0, 0.0))
        wr, dr = right.get(c, (0.0, 0.0))
        np.testing.assert_allclose(wl, wr, rtol=0, atol=1e-6 * np.abs(left[8][0]).max())
        np.testing.assert_allclose(dl, dr, rtol=0, atol=1e-6 * scale)


def small_rollout(n_nodes=40, grid=16, seed=3):
    shape = (grid, grid, grid)
    omega0 = synthetic.random_phase_vorticity(shape, 1.0, max_wavenumber=3, seed=seed,
                                              max_vorticity=4.0)
    X0 = synthetic.fibonacci_ellipsoid(n_nodes, (0.2, 0.12, 0.12), (0.5, 0.5, 0.5))
    kin = A.HeaveKinematics(X0, period=0.02)
    return A.Rollout(shape, 1.0, 1e-2, omega0, (1.0, 0.0, 0.0), kin, k=-60.0, c=-3.0)


def test_rollout_gradient_matches_finite_differences():
    ro = small_rollout()
    rng = np.random.default_rng(7)
    theta = rng.normal(0.0, 0.5, 16)
    nsteps = 6
    _, dts = ro.forward(theta, nsteps)
    wts = A.time_average_weights(dts)
    fbar = np.zeros((nsteps, 3))
    fbar[:, 0] = wts

    def J(th):
        f, _ = ro.forward(th, nsteps, dts=dts)
        return float(f[:, 0] @ wts)

    g = ro.gradient(theta, nsteps, dts, fbar)
    h = 1e-5
    fd = np.array([(J(theta + h * e) - J(theta - h * e)) / (2 * h) for e in np.eye(16)])
    err = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    assert err < 1e-6, (err, g, fd)
    # the node-position path (dw/dX) carries a measurable part of the gradient
    gv = ro.gradient(theta, nsteps, dts, fbar, include_positions=False)
    assert np.linalg.norm(gv - fd) / np.linalg.norm(fd) > 100 * err
