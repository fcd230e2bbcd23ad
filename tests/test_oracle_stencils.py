"""Pins the oracle's grid stencils with closed forms and invariants (CPU only).

Each check is chosen so a plausible slip (a sign, a swapped axis, a dropped term, a wrong
write region) fails: fields are linear/quadratic in x, y, z with distinct coefficients per
component and axis, so every finite difference is exact and identifies its axis.
"""

import numpy as np
import pytest

from oracle import stencils as S


def coords(n, dx=1.0):
    z, y, x = np.meshgrid(*(np.arange(n, dtype=np.float64) * dx,) * 3, indexing="ij")
    return x, y, z


def test_curl_linear_field_exact_and_border():
    n = 9
    x, y, z = coords(n)
    # f = (a.x, b.x, c.x) with distinct coefficient vectors a, b, c
    a, b, c = np.array([1.0, 2.0, 3.0]), np.array([5.0, 7.0, 11.0]), np.array([13.0, 17.0, 19.0])
    f = np.stack([a[0] * x + a[1] * y + a[2] * z, b[0] * x + b[1] * y + b[2] * z, c[0] * x + c[1] * y + c[2] * z])
    out = np.ones_like(f)
    S.curl(out, f, 0.5)  # prefactor 0.5 turns central sums into derivatives (dx = 1)
    # curl = (dfz/dy - dfy/dz, dfx/dz - dfz/dx, dfy/dx - dfx/dy)
    expect = np.array([c[1] - b[2], a[2] - c[0], b[0] - a[1]])
    for i in range(3):
        np.testing.assert_allclose(out[i][1:-1, 1:-1, 1:-1], expect[i], rtol=0, atol=1e-12)
        border = out[i].copy()
        border[1:-1, 1:-1, 1:-1] = 0
        assert np.all(border == 0)
    out2 = np.full_like(f, 7.0)
    S.curl(out2, f, 0.5, reset_ghost_zone=False)
    assert np.all(out2[:, 0] == 7.0) and np.all(out2[:, :, :, -1] == 7.0)


def test_div_curl_vanishes():
    rng = np.random.default_rng(0)
    psi = rng.standard_normal((3, 10, 11, 12))
    u = np.zeros_like(psi)
    S.curl(u, psi, 0.5)
    div = np.zeros(psi.shape[1:])
    S.divergence(div, u, 1.0)
    np.testing.assert_allclose(div[2:-2, 2:-2, 2:-2], 0.0, atol=1e-12)


def test_update_vorticity_adds_curl_interior_only():
    rng = np.random.default_rng(1)
    f = rng.standard_normal((3, 8, 8, 8))
    w = rng.standard_normal((3, 8, 8, 8))
    w0 = w.copy()
    S.update_vorticity_from_velocity_forcing(w, f, 0.3)
    c = np.zeros_like(f)
    S.curl(c, f, 0.3)
    np.testing.assert_allclose(w - w0, c, atol=1e-14)
    # penalised variant with u_pen - u == f
    u = rng.standard_normal(f.shape)
    w1 = w0.copy()
    S.update_vorticity_from_penalised_velocity(w1, u + f, u, 0.3)
    np.testing.assert_allclose(w1, w, atol=1e-12)


def test_diffusion_quadratic_field():
    n = 8
    x, y, z = coords(n)
    f = 1.5 * x**2 - 2.0 * y**2 + 4.0 * z**2 + 3 * x * y
    flux = np.ones_like(f)
    S.diffusion_flux(flux, f, 0.25)
    # discrete laplacian of a quadratic is exact: 2*(1.5 - 2 + 4) = 7
    np.testing.assert_allclose(flux[1:-1, 1:-1, 1:-1], 0.25 * 7.0, atol=1e-12)
    b = flux.copy()
    b[1:-1, 1:-1, 1:-1] = 0
    assert np.all(b == 0)
    v = np.stack([f, 2 * f, -f])
    v0 = v.copy()
    buf = np.zeros_like(f)
    S.diffusion_timestep_euler_forward(v, buf, 0.25)
    for i, s in enumerate([1, 2, -1]):
        np.testing.assert_allclose(v[i][1:-1, 1:-1, 1:-1] - v0[i][1:-1, 1:-1, 1:-1], s * 1.75, atol=1e-12)
        np.testing.assert_array_equal(v[i][0], v0[i][0])


def test_cross_product_closed_form():
    # test_elementwise_ops_3d.py:286-309: [1,2,3] x [4,5,6] = [-3, 6, -3]
    a = np.stack([np.full((4, 4, 4), v) for v in (1.0, 2.0, 3.0)])
    b = np.stack([np.full((4, 4, 4), v) for v in (4.0, 5.0, 6.0)])
    r = np.zeros_like(a)
    S.elementwise_cross_product(r, a, b)
    for i, e in enumerate((-3.0, 6.0, -3.0)):
        assert np.all(r[i] == e)


def _penalise_closed_form(f, w, dx, xs, ys, zs):
    """Independent formulation: value at the clamped index times the three slab factors."""
    nz, ny, nx = f.shape
    sp = (np.pi / 2) / (w * dx)

    def fac(c):
        n = len(c)
        s = np.ones(n)
        s[:w] = np.sin(sp * (c[:w] - c[0]))
        s[n - w :] = np.sin(sp * (c[-1] - c[n - w :]))
        return s

    ci = lambda n: np.clip(np.arange(n), w - 1, n - w)  # noqa: E731
    out = f[np.ix_(ci(nz), ci(ny), ci(nx))]
    return ((out * fac(xs)[None, None, :]) * fac(ys)[None, :, None]) * fac(zs)[:, None, None]


@pytest.mark.parametrize("width", [1, 2, 4])
def test_penalisation_two_formulations_agree(width):
    rng = np.random.default_rng(2)
    n = 16
    dx = 0.1
    c = np.linspace(dx / 2, 1 - dx / 2, n)
    zg, yg, xg = np.meshgrid(c, c, c, indexing="ij")
    f = rng.standard_normal((n, n, n))
    g = f.copy()
    S.penalise_field_boundary(g, width, dx, xg, yg, zg)
    np.testing.assert_allclose(g, _penalise_closed_form(f, width, dx, c, c, c), rtol=0, atol=1e-14)
    # the interior is untouched; the outermost planes are exactly zero
    np.testing.assert_array_equal(g[width:-width, width:-width, width:-width], f[width:-width, width:-width, width:-width])
    assert np.all(g[0] == 0) and np.all(g[:, :, -1] == 0)


def test_penalisation_width_zero_noop():
    f = np.random.default_rng(3).standard_normal((6, 6, 6))
    g = f.copy()
    S.penalise_field_boundary(g, 0, 0.1, f, f, f)
    np.testing.assert_array_equal(f, g)


def test_set_add_boundaries_saxpby():
    f = np.zeros((3, 5, 6, 7))
    S.set_fixed_val(f, (1.0, 2.0, 3.0))
    assert np.all(f[1] == 2.0)
    S.add_fixed_val(f, f, (1.0, 1.0, -3.0))
    assert np.all(f[2] == 0.0) and np.all(f[0] == 2.0)
    s = np.ones((5, 6, 7))
    S.set_fixed_val_at_boundaries(s, 2, 0.0)
    assert s.sum() == (5 - 4) * (6 - 4) * (7 - 4)
    r = np.zeros((5, 6, 7))
    S.elementwise_saxpby(r, np.full_like(r, 2.0), np.full_like(r, 3.0), 0.5, -1.0)
    assert np.all(r == -2.0)


def test_eno3_constant_velocity_linear_field():
    # conservative flux of u f with u constant > 0 and f linear in x: the upwind-biased
    # third-order difference of u f is exact for linear f, so flux = -u * df/dx * dt/dx * dx
    n = 10
    x, y, z = coords(n)
    f = 3.0 * x + 0.0 * y
    u = np.stack([np.full(f.shape, 2.0), np.zeros(f.shape), np.zeros(f.shape)])
    flux = np.zeros_like(f)
    S.advection_flux_conservative_eno3(flux, f, u, 1.0)
    # inv_dx * (F(i+1/2) - F(i-1/2)) with F the face flux of u f; d(uf)/dx = 6
    np.testing.assert_allclose(flux[2:-2, 2:-2, 2:-2], 6.0, atol=1e-12)
    b = flux.copy()
    b[2:-2, 2:-2, 2:-2] = 0
    assert np.all(b == 0)
    # negative velocity takes the other branch: still exact
    flux[...] = 0
    S.advection_flux_conservative_eno3(flux, f, -u, 1.0)
    np.testing.assert_allclose(flux[2:-2, 2:-2, 2:-2], -6.0, atol=1e-12)


def _eno3_masks(u):
    n = u.shape[1:]
    J = (slice(2, n[0] - 2), slice(2, n[1] - 2), slice(2, n[2] - 2))
    out = []
    for axis, comp in ((2, 0), (1, 1), (0, 2)):
        for off in (1, -1):
            sl = list(J)
            sl[axis] = slice(J[axis].start + off, J[axis].stop + off)
            out.append(u[comp][J] > -u[comp][tuple(sl)])
    return np.stack(out)


def test_eno3_adjoint_dot_products():
    # the forward is linear in the field and, with the switches frozen, in the velocity, so
    # <F(f; u), a> = <f, J_f^T a> = <u, J_u^T a>, and a small velocity change that flips no
    # switch moves <F, a> by exactly <du, J_u^T a>
    rng = np.random.default_rng(7)
    shape = (8, 9, 10)
    f = rng.standard_normal(shape)
    u = rng.standard_normal((3, *shape))
    a = rng.standard_normal(shape)
    inv_dx = 0.7
    flux = np.zeros(shape)
    S.advection_flux_conservative_eno3(flux, f, u, inv_dx)
    fa = np.zeros(shape)
    ua = np.zeros((3, *shape))
    S.advection_flux_conservative_eno3_adjoint(fa, ua, a, f, u, inv_dx)
    lhs = np.vdot(flux, a)
    assert abs(lhs) > 0.1
    np.testing.assert_allclose(np.vdot(f, fa), lhs, rtol=1e-12)
    np.testing.assert_allclose(np.vdot(u, ua), lhs, rtol=1e-12)
    du = 1e-7 * rng.standard_normal(u.shape)
    assert np.array_equal(_eno3_masks(u), _eno3_masks(u + du))
    flux2 = np.zeros(shape)
    S.advection_flux_conservative_eno3(flux2, f, u + du, inv_dx)
    np.testing.assert_allclose(np.vdot(flux2 - flux, a), np.vdot(du, ua), rtol=1e-6)
    # each unit field adjoint vector: a column of the transpose equals a row of the forward
    for c in [(2, 2, 2), (5, 4, 7), (3, 6, 2)]:
        e = np.zeros(shape)
        e[c] = 1.0
        fa = np.zeros(shape)
        S.advection_flux_conservative_eno3_adjoint(fa, None, e, f, u, inv_dx)
        for m in [(c[0], c[1], c[2] - 2), (c[0], c[1] + 1, c[2]), (c[0] - 1, c[1], c[2]), c]:
            g = np.zeros(shape)
            g[m] = 1.0
            fl = np.zeros(shape)
            S.advection_flux_conservative_eno3(fl, g, u, inv_dx)
            np.testing.assert_allclose(fa[m], fl[c], rtol=1e-12, atol=1e-15)


def test_vorticity_stretching_linear_velocity():
    n = 7
    x, y, z = coords(n)
    w = np.stack([np.full(x.shape, 1.0), np.full(x.shape, 2.0), np.full(x.shape, 3.0)])
    # u_c = g_c . x, so (w . grad) u_c = w . g_c, and the unscaled central difference gives 2 w . g_c
    g = np.array([[1.0, 0.5, -2.0], [0.0, 3.0, 1.0], [4.0, -1.0, 0.5]])
    u = np.stack([g[c, 0] * x + g[c, 1] * y + g[c, 2] * z for c in range(3)])
    flux = np.ones_like(u)
    S.vorticity_stretching_flux(flux, w, u, 0.5)
    for c in range(3):
        np.testing.assert_allclose(flux[c][1:-1, 1:-1, 1:-1], np.dot(g[c], [1.0, 2.0, 3.0]), atol=1e-12)
        assert np.all(flux[c][0] == 0)


# ---------------------------------------------------------------- Laplacian filter (A7)
# Pins of the oracle's laplacian_filter (laplacian_filter_3d.py:13-194):
# * a constant field is unchanged (the reference's test_laplacian_filter_constant_field,
#   test_laplacian_filter_3d.py:97-131);
# * a separable cosine mode cos(a x) cos(b y) cos(c z) is an eigenfunction of each 1D pass
#   away from the borders, with factor s_i = (1 - cos(k_i h)) / 2: far from the border the
#   multiplicative filter of order k leaves (1 - (s_x s_y s_z)^k) f and the convolution
#   filter (1 - s_x^k)(1 - s_y^k)(1 - s_z^k) f (a sign, coefficient, axis or order mistake
#   changes these numbers);
# * border cells are never changed (the flux is zero there).


def _mode(n, kx, ky, kz):
    i = np.arange(n)
    ax, ay, az = 2 * np.pi * kx / n, 2 * np.pi * ky / n, 2 * np.pi * kz / n
    f = np.cos(az * i)[:, None, None] * np.cos(ay * i)[None, :, None] * np.cos(ax * i)[None, None, :]
    s = [(1 - np.cos(a)) / 2 for a in (ax, ay, az)]
    return f, s


@pytest.mark.parametrize("filter_type", ["multiplicative", "convolution"])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_laplacian_filter_cosine_mode_closed_form(filter_type, order):
    n = 32
    f, (sx, sy, sz) = _mode(n, 3, 5, 2)
    g = f.copy()
    S.laplacian_filter(g, order, filter_type, np.zeros_like(f), np.zeros_like(f))
    if filter_type == "multiplicative":
        factor = 1 - (sx * sy * sz) ** order
    else:
        factor = (1 - sx**order) * (1 - sy**order) * (1 - sz**order)
    m = 3 * order + 1  # the border's influence travels one cell per pass
    c = (slice(m, -m),) * 3
    np.testing.assert_allclose(g[c], factor * f[c], rtol=0, atol=1e-12)
    for sl in S._boundary_slabs(1):
        np.testing.assert_array_equal(g[sl], f[sl])


@pytest.mark.parametrize("filter_type", ["multiplicative", "convolution"])
@pytest.mark.parametrize("order", [0, 1, 2])
def test_laplacian_filter_constant_field_unchanged(filter_type, order):
    f = 2 * np.ones((16, 16, 16))
    g = f.copy()
    S.laplacian_filter(g, order, filter_type, np.zeros_like(f), np.zeros_like(f))
    np.testing.assert_allclose(g, f, atol=1e-14)
