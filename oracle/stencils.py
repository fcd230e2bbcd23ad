"""Oracle: 3D grid stencils of the flow step (numpy restatement, test infrastructure only).

Layout: vector fields are ``(3, Nz, Ny, Nx)``, components ordered x, y, z
(``sopht/utils/field.py:8-39``); scalar fields are ``(Nz, Ny, Nx)``. Array axis 0 of a
scalar field is z, axis 2 is x (x fastest).

Write-region convention (library behaviour of pystencils, pinned by the reference's own
whole-array tests): a neighbour stencil writes only the interior whose ghost width equals
the largest absolute offset of the stencil, on every axis. Radius-1 stencils write
``[1:-1]^3``; ENO3 stencils write ``[2:-2]^3``; pointwise kernels write everything.

Pinning (tests/test_oracle_stencils.py): linear/quadratic fields with exact finite
differences, ``div_h(curl_h psi) == 0``, the constant-field closed forms of
``tests/test_numeric/test_eulerian_grid_ops/test_stencil_ops_3d/test_elementwise_ops_3d.py``,
and a second independent formulation of the penalisation (clamped-index closed form).
"""

from __future__ import annotations

import numpy as np

I = slice(1, -1)  # radius-1 interior along one axis


def _d(a, axis, off):
    """Interior slice of ``a`` shifted by ``off`` along spatial axis ``axis`` (0=z,1=y,2=x)."""
    sl = [I, I, I]
    n = a.shape[axis]
    sl[axis] = slice(1 + off, n - 1 + off)
    return a[tuple(sl)]


def curl_h(field: np.ndarray, prefactor) -> tuple:
    """Interior of ``prefactor * curl_h(field)`` per component.

    Follows ``curl_3d.py:37-39,50-52,64-66``:
    curl_x = f_z[y+1]-f_z[y-1]-f_y[z+1]+f_y[z-1]; curl_y = f_x[z+1]-f_x[z-1]-f_z[x+1]+f_z[x-1];
    curl_z = f_y[x+1]-f_y[x-1]-f_x[y+1]+f_x[y-1].
    """
    fx, fy, fz = field[0], field[1], field[2]
    cx = prefactor * (_d(fz, 1, 1) - _d(fz, 1, -1) - _d(fy, 0, 1) + _d(fy, 0, -1))
    cy = prefactor * (_d(fx, 0, 1) - _d(fx, 0, -1) - _d(fz, 2, 1) + _d(fz, 2, -1))
    cz = prefactor * (_d(fy, 2, 1) - _d(fy, 2, -1) - _d(fx, 1, 1) + _d(fx, 1, -1))
    return cx, cy, cz


def curl(curl_out: np.ndarray, field: np.ndarray, prefactor, reset_ghost_zone=True) -> None:
    """``gen_curl_pyst_kernel_3d`` (``curl_3d.py:13-132``): interior curl; border set to 0
    when ``reset_ghost_zone`` (``:103-130``), else left untouched."""
    cx, cy, cz = curl_h(field, prefactor)
    curl_out[0][I, I, I] = cx
    curl_out[1][I, I, I] = cy
    curl_out[2][I, I, I] = cz
    if reset_ghost_zone:
        set_fixed_val_at_boundaries(curl_out, 1, (0, 0, 0))


def update_vorticity_from_velocity_forcing(vorticity_field, velocity_forcing_field, prefactor):
    """``update_vorticity_from_velocity_forcing_3d.py:12-132``: ``w[1:-1]^3 += p*curl_h(f)``;
    the border of ``w`` is untouched. Components are updated x, y, z in turn, each reading
    only ``f`` (so the order does not matter)."""
    cx, cy, cz = curl_h(velocity_forcing_field, prefactor)
    vorticity_field[0][I, I, I] += cx
    vorticity_field[1][I, I, I] += cy
    vorticity_field[2][I, I, I] += cz


def update_vorticity_from_penalised_velocity(
    vorticity_field, penalised_velocity_field, velocity_field, prefactor
):
    """``update_vorticity_from_velocity_forcing_3d.py:135-291``: ``w += p*curl_h(u_pen - u)``
    on the interior (the reference expands the difference term by term)."""
    cx, cy, cz = curl_h(penalised_velocity_field - velocity_field, prefactor)
    vorticity_field[0][I, I, I] += cx
    vorticity_field[1][I, I, I] += cy
    vorticity_field[2][I, I, I] += cz


def elementwise_cross_product(result_field, field_1, field_2):
    """``elementwise_ops_3d.py:390-449``: ``r_i = a_j b_k - b_j a_k`` with cyclic (i, j, k)
    over all cells."""
    a, b = field_1, field_2
    result_field[0] = a[1] * b[2] - b[1] * a[2]
    result_field[1] = a[2] * b[0] - b[2] * a[0]
    result_field[2] = a[0] * b[1] - b[0] * a[1]


def diffusion_flux(diffusion_flux_field, field, prefactor, reset_ghost_zone=True):
    """``diffusion_flux_3d.py:14-75`` (scalar): ``flux[1:-1]^3 = p*(sum of 6 nbrs - 6c)``;
    flux border 0 when ``reset_ghost_zone``."""
    f = field
    lap = (
        _d(f, 0, 1)
        + _d(f, 0, -1)
        + _d(f, 1, 1)
        + _d(f, 1, -1)
        + _d(f, 2, 1)
        + _d(f, 2, -1)
        - 6 * f[I, I, I]
    )
    diffusion_flux_field[I, I, I] = prefactor * lap
    if reset_ghost_zone:
        set_fixed_val_at_boundaries(diffusion_flux_field, 1, 0)


def diffusion_timestep_euler_forward(field, diffusion_flux_buffer, nu_dt_by_dx2):
    """``diffusion_timestep_3d.py:12-80``: per scalar (or per component of a vector, in
    order x, y, z, reusing one scalar flux buffer): flux with border reset, then
    ``field += flux`` on all cells (so the border of ``field`` is unchanged)."""
    comps = [field] if field.ndim == 3 else [field[0], field[1], field[2]]
    for c in comps:
        diffusion_flux(diffusion_flux_buffer, c, nu_dt_by_dx2, reset_ghost_zone=True)
        c[...] = c + diffusion_flux_buffer


def divergence(divergence_field, field, inv_dx, reset_ghost_zone=True):
    """``divergence_3d.py:13-96``: ``div[1:-1]^3 = 0.5*inv_dx*(dfx/dx + dfy/dy + dfz/dz)``
    with central differences; border 0 when ``reset_ghost_zone``."""
    fx, fy, fz = field[0], field[1], field[2]
    divergence_field[I, I, I] = (
        0.5
        * inv_dx
        * (
            _d(fx, 2, 1)
            - _d(fx, 2, -1)
            + _d(fy, 1, 1)
            - _d(fy, 1, -1)
            + _d(fz, 0, 1)
            - _d(fz, 0, -1)
        )
    )
    if reset_ghost_zone:
        set_fixed_val_at_boundaries(divergence_field, 1, 0)


def set_fixed_val(field, fixed_val):
    """``elementwise_ops_3d.py:60-119``: scalar fill, or per-component fill for a vector."""
    if field.ndim == 3:
        field[...] = fixed_val
    else:
        for i in range(3):
            field[i] = fixed_val[i]


def add_fixed_val(sum_field, field, fixed_val):
    """``elementwise_ops_3d.py:271-334``: ``sum = field + val`` (per component for vectors)."""
    if field.ndim == 3:
        sum_field[...] = field + fixed_val
    else:
        for i in range(3):
            sum_field[i] = field[i] + fixed_val[i]


def elementwise_sum(sum_field, field_1, field_2):
    """``elementwise_ops_3d.py:13-57``."""
    sum_field[...] = field_1 + field_2


def elementwise_copy(field, rhs_field):
    """``elementwise_ops_3d.py:122-143``."""
    field[...] = rhs_field


def elementwise_saxpby(sum_field, field_1, field_2, field_1_prefac, field_2_prefac):
    """``elementwise_ops_3d.py:337-387``: ``s = a*x + b*y``."""
    sum_field[...] = field_1_prefac * field_1 + field_2_prefac * field_2


def elementwise_complex_product(product_field, field_1, field_2):
    """``elementwise_ops_3d.py:146-197``: complex product written as real/imag parts."""
    re = field_1.real * field_2.real - field_1.imag * field_2.imag
    im = field_1.real * field_2.imag + field_1.imag * field_2.real
    product_field.real = re
    product_field.imag = im


def set_fixed_val_at_boundaries(field, width, fixed_val):
    """``elementwise_ops_3d.py:200-268``: fill slabs of ``width`` on the 6 faces."""
    if field.ndim == 3:
        for sl in _boundary_slabs(width):
            field[sl] = fixed_val
    else:
        for i in range(3):
            for sl in _boundary_slabs(width):
                field[i][sl] = fixed_val[i]


def _boundary_slabs(w):
    a = slice(None)
    return [
        (slice(None, w), a, a),
        (slice(-w, None), a, a),
        (a, slice(None, w), a),
        (a, slice(-w, None), a),
        (a, a, slice(None, w)),
        (a, a, slice(-w, None)),
    ]


def penalise_field_boundary(field, width, dx, x_grid_field, y_grid_field, z_grid_field):
    """``penalise_field_boundary_3d.py:13-240`` on a scalar field, in place.

    Sequence of ``:191-208``: along x then y then z, copy the plane at index ``w-1``
    (resp. ``-w``) into the ``w``-wide border slab, then multiply the slab by
    ``sin(pi/(2 w dx) * distance from the first/last cell centre)``. ``width == 0`` is a
    no-op (``:29-50``).

    The back-slab source is the plane ``n - w`` written with explicit indices: the
    reference's ``-w:(-w+1)`` slice is empty for ``w == 1`` and raises there; for every
    other width the two spellings are the same plane.
    """
    if width == 0:
        return
    w = width
    sp = (np.pi / 2) / (w * dx)
    x0, xe = x_grid_field[0, 0, 0], x_grid_field[0, 0, -1]
    y0, ye = y_grid_field[0, 0, 0], y_grid_field[0, -1, 0]
    z0, ze = z_grid_field[0, 0, 0], z_grid_field[-1, 0, 0]
    f = field
    nz, ny, nx = f.shape
    f[:, :, :w] = f[:, :, (w - 1) : w]
    f[:, :, -w:] = f[:, :, nx - w : nx - w + 1]
    f[:, :, :w] *= np.sin(sp * (x_grid_field[:, :, :w] - x0))
    f[:, :, -w:] *= np.sin(sp * (xe - x_grid_field[:, :, -w:]))
    f[:, :w, :] = f[:, (w - 1) : w, :]
    f[:, -w:, :] = f[:, ny - w : ny - w + 1, :]
    f[:, :w, :] *= np.sin(sp * (y_grid_field[:, :w, :] - y0))
    f[:, -w:, :] *= np.sin(sp * (ye - y_grid_field[:, -w:, :]))
    f[:w, :, :] = f[(w - 1) : w, :, :]
    f[-w:, :, :] = f[nz - w : nz - w + 1, :, :]
    f[:w, :, :] *= np.sin(sp * (z_grid_field[:w, :, :] - z0))
    f[-w:, :, :] *= np.sin(sp * (ze - z_grid_field[-w:, :, :]))


def penalise_vector_field_boundary(vector_field, width, dx, x_grid, y_grid, z_grid):
    """Vector form (``penalise_field_boundary_3d.py:210-237``): each component in turn."""
    for i in range(3):
        penalise_field_boundary(vector_field[i], width, dx, x_grid, y_grid, z_grid)


# ---------------------------------------------------------------- off-step factories (A23, A24)


def advection_flux_conservative_eno3(advection_flux, field, velocity, inv_dx):
    """``advection_flux_3d.py:12-233``: six sequential sub-kernels, each accumulating into
    ``advection_flux`` on ``[2:-2]^3`` (ghost width 2 on every axis). Order: x front
    (``+``), x back (``-``), y front, y back, z front, z back. Upwind switch of the front
    face: ``u[0] > -u[+1]``; of the back face: ``u[0] > -u[-1]``."""
    n = field.shape
    J = (slice(2, n[0] - 2), slice(2, n[1] - 2), slice(2, n[2] - 2))

    def sh(a, axis, off):
        sl = list(J)
        s = sl[axis]
        sl[axis] = slice(s.start + off, s.stop + off)
        return a[tuple(sl)]

    for axis, comp in ((2, 0), (1, 1), (0, 2)):
        u = velocity[comp]
        f = field
        # front face
        up = sh(u, axis, 0) > -sh(u, axis, 1)
        a = (1 / 3) * sh(f, axis, 1) * sh(u, axis, 1) + (5 / 6) * sh(f, axis, 0) * sh(u, axis, 0) - (
            1 / 6
        ) * sh(f, axis, -1) * sh(u, axis, -1)
        b = (1 / 3) * sh(f, axis, 0) * sh(u, axis, 0) + (5 / 6) * sh(f, axis, 1) * sh(u, axis, 1) - (
            1 / 6
        ) * sh(f, axis, 2) * sh(u, axis, 2)
        advection_flux[J] = advection_flux[J] + inv_dx * np.where(up, a, b)
        # back face
        up = sh(u, axis, 0) > -sh(u, axis, -1)
        a = (1 / 3) * sh(f, axis, 0) * sh(u, axis, 0) + (5 / 6) * sh(f, axis, -1) * sh(u, axis, -1) - (
            1 / 6
        ) * sh(f, axis, -2) * sh(u, axis, -2)
        b = (1 / 3) * sh(f, axis, -1) * sh(u, axis, -1) + (5 / 6) * sh(f, axis, 0) * sh(u, axis, 0) - (
            1 / 6
        ) * sh(f, axis, 1) * sh(u, axis, 1)
        advection_flux[J] = advection_flux[J] - inv_dx * np.where(up, a, b)


def advection_flux_conservative_eno3_adjoint(field_adj, velocity_adj, flux_adj, field, velocity, inv_dx):
    """Transpose of ``advection_flux_conservative_eno3`` (``advection_flux_3d.py:12-233``) with
    each face's stencil choice frozen at the forward's ``velocity``. The reference has no
    adjoint; this restates the forward's transpose in scatter form.

    The forward adds ``inv_dx * sum_o W(c, o) * g[c + o]`` to the flux at ``c`` in
    ``[2:-2]^3``, with ``g = field * u_axis`` and the face weights ``W`` picked by the
    upwind switches (which read only ``velocity``). So for a flux adjoint ``a``,
    ``gbar[c + o] += inv_dx * W(c, o) * a[c]``, then ``field_adj += gbar * u_axis`` and
    ``velocity_adj[axis] += gbar * field``. Either output may be None."""
    n = field.shape
    J = (slice(2, n[0] - 2), slice(2, n[1] - 2), slice(2, n[2] - 2))

    def sh(a, axis, off):
        sl = list(J)
        s = sl[axis]
        sl[axis] = slice(s.start + off, s.stop + off)
        return a[tuple(sl)]

    a = flux_adj[J]
    for axis, comp in ((2, 0), (1, 1), (0, 2)):
        u = velocity[comp]
        gbar = np.zeros(n, dtype=np.result_type(field, flux_adj))
        upf = sh(u, axis, 0) > -sh(u, axis, 1)
        upb = sh(u, axis, 0) > -sh(u, axis, -1)
        # (offset, weight when the switch holds, weight when it does not); back faces negated
        terms = (
            (upf, ((1, 1 / 3), (0, 5 / 6), (-1, -1 / 6)), ((0, 1 / 3), (1, 5 / 6), (2, -1 / 6)), 1.0),
            (upb, ((0, 1 / 3), (-1, 5 / 6), (-2, -1 / 6)), ((-1, 1 / 3), (0, 5 / 6), (1, -1 / 6)), -1.0),
        )
        for mask, taken, other, sign in terms:
            for off, w in taken:
                sh(gbar, axis, off)[...] += np.where(mask, sign * w * inv_dx * a, 0)
            for off, w in other:
                sh(gbar, axis, off)[...] += np.where(mask, 0, sign * w * inv_dx * a)
        if field_adj is not None:
            field_adj += gbar * u
        if velocity_adj is not None:
            velocity_adj[comp] += gbar * field


def advection_timestep_euler_forward_conservative_eno3(field, advection_flux, velocity, dt_by_dx):
    """``advection_timestep_3d.py:12-99``: per scalar (or per component, x, y, z): zero the
    flux, accumulate the ENO3 flux with ``inv_dx = -dt_by_dx``, then ``field += flux``."""
    comps = [field] if field.ndim == 3 else [field[0], field[1], field[2]]
    for c in comps:
        advection_flux[...] = 0
        advection_flux_conservative_eno3(advection_flux, c, velocity, -dt_by_dx)
        c[...] = c + advection_flux


def vorticity_stretching_flux(flux_field, vorticity_field, velocity_field, prefactor):
    """``vorticity_stretching_flux_3d.py:13-107``: ``flux_c[1:-1]^3 =
    p*(w_x du_c/dx + w_y du_c/dy + w_z du_c/dz)`` (central, unscaled differences); border 0."""
    wx, wy, wz = (vorticity_field[i][I, I, I] for i in range(3))
    for c in range(3):
        u = velocity_field[c]
        flux_field[c][I, I, I] = prefactor * (
            wx * (_d(u, 2, 1) - _d(u, 2, -1))
            + wy * (_d(u, 1, 1) - _d(u, 1, -1))
            + wz * (_d(u, 0, 1) - _d(u, 0, -1))
        )
    set_fixed_val_at_boundaries(flux_field, 1, (0, 0, 0))


def vorticity_stretching_timestep_euler_forward(vorticity_field, velocity_field, flux_field, dt_by_2_dx):
    """``vorticity_stretching_timestep_3d.py:10-55``."""
    vorticity_stretching_flux(flux_field, vorticity_field, velocity_field, dt_by_2_dx)
    vorticity_field[...] = vorticity_field + flux_field


def vorticity_stretching_timestep_ssprk3(
    vorticity_field, velocity_field, flux_field, dt_by_2_dx, midstep_buffer
):
    """``vorticity_stretching_timestep_3d.py:58-158``. Both mid-step views alias one buffer
    (``:113,125``); the third stage uses prefactor ``dt_by_2_dx * 0.5`` (``:137``)."""
    m = midstep_buffer
    vorticity_stretching_flux(flux_field, vorticity_field, velocity_field, dt_by_2_dx)
    m[...] = vorticity_field + flux_field
    vorticity_stretching_flux(flux_field, m, velocity_field, dt_by_2_dx)
    m[...] = m + flux_field
    m[...] = 0.75 * vorticity_field + 0.25 * m
    vorticity_stretching_flux(flux_field, m, velocity_field, dt_by_2_dx * 0.5)
    m[...] = m + flux_field
    vorticity_field[...] = (1.0 / 3.0) * vorticity_field + (2.0 / 3.0) * m


def laplacian_filter_1d(filter_flux, field, axis):
    """One 1D Laplacian filter pass (``laplacian_filter_3d.py:60-80``): on ``[1:-1]^3``,
    ``flux = 0.25 * (-f[+1] - f[-1] + 2 f)`` along ``axis`` (0 = z, 1 = y, 2 = x); the rest of
    ``filter_flux`` is left as it is."""
    c = field[I, I, I]
    filter_flux[I, I, I] = 0.25 * ((-_d(field, axis, 1) - _d(field, axis, -1)) + 2 * c)


def laplacian_filter(scalar_field, filter_order, filter_type, filter_flux_buffer, field_buffer,
                     boundary_width=1):
    """``gen_laplacian_filter_kernel_3d`` on a scalar field (``laplacian_filter_3d.py:98-166``),
    in place. The flux buffer's boundary of ``boundary_width`` is zeroed first; each pass writes
    the interior and the field buffer becomes a full copy of the flux.
    multiplicative: ``order`` times x, y, z passes, then ``field -= flux``;
    convolution: per axis x, y, z: ``order`` passes from a copy of the field, then ``field -= flux``."""
    set_fixed_val_at_boundaries(filter_flux_buffer, boundary_width, 0)
    axes = (2, 1, 0)  # x, y, z
    if filter_type == "multiplicative":
        field_buffer[...] = scalar_field
        for _ in range(filter_order):
            for a in axes:
                laplacian_filter_1d(filter_flux_buffer, field_buffer, a)
                field_buffer[...] = filter_flux_buffer
        scalar_field[...] = scalar_field - filter_flux_buffer
    elif filter_type == "convolution":
        for a in axes:
            field_buffer[...] = scalar_field
            for _ in range(filter_order):
                laplacian_filter_1d(filter_flux_buffer, field_buffer, a)
                field_buffer[...] = filter_flux_buffer
            scalar_field[...] = scalar_field - filter_flux_buffer
    else:
        raise ValueError("Invalid filter type")


def laplacian_filter_vector(vector_field, filter_order, filter_type, filter_flux_buffer, field_buffer):
    """Vector form (``laplacian_filter_3d.py:176-189``): components x, y, z in turn."""
    for i in range(3):
        laplacian_filter(vector_field[i], filter_order, filter_type, filter_flux_buffer, field_buffer)
