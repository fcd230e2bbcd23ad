"""Drop-in ``gen_*_pyst_kernel_3d`` factories backed by the gfx950 HIP library.

Same factory signatures, keyword names and write semantics as
``sopht/numeric/eulerian_grid_ops/stencil_ops_3d/*.py``; operands are torch tensors on the
HIP device (strided views accepted), launched on torch's current stream. ``num_threads``
and ``fixed_grid_size`` are accepted and ignored, as in the reference (its
``fixed_grid_size`` check never matches, e.g. ``curl_3d.py:24-28``).
"""

from __future__ import annotations

from collections.abc import Callable

import numpy as np

from ... import _lib as L


def _check_field_type(field_type):
    if field_type not in ("scalar", "vector"):
        msg = "Invalid field type"
        raise ValueError(msg)


def _shape3(t):
    return tuple(int(s) for s in t.shape[-3:])


# ------------------------------------------------------------------------------ curl family


def gen_curl_pyst_kernel_3d(
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
    reset_ghost_zone: bool = True,
) -> Callable:
    """``curl_3d.py:13-132``: ``curl[1:-1]^3 = prefactor * curl_h(field)``; border 0 when
    ``reset_ghost_zone``."""
    tag = L.dtype_tag(real_t)

    def curl_pyst_kernel_3d(curl, field, prefactor: float) -> None:
        nz, ny, nx = _shape3(field)
        L.call(
            "sopht_curl_3d", tag, L.comps(curl), L.strides3(curl[0]), L.comps(field),
            L.strides3(field[0]), nz, ny, nx, float(prefactor), int(bool(reset_ghost_zone)),
            L.stream_ptr(),
        )
        L.written(curl)

    return curl_pyst_kernel_3d


def gen_update_vorticity_from_velocity_forcing_pyst_kernel_3d(
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
) -> Callable:
    """``update_vorticity_from_velocity_forcing_3d.py:12-132``:
    ``vorticity[1:-1]^3 += prefactor * curl_h(velocity_forcing)``."""
    tag = L.dtype_tag(real_t)

    def update_vorticity_from_velocity_forcing_pyst_kernel_3d(
        vorticity_field, velocity_forcing_field, prefactor: float
    ) -> None:
        nz, ny, nx = _shape3(vorticity_field)
        L.call(
            "sopht_update_vorticity_from_velocity_forcing_3d", tag, L.comps(vorticity_field),
            L.strides3(vorticity_field[0]), L.comps(velocity_forcing_field),
            L.strides3(velocity_forcing_field[0]), nz, ny, nx, float(prefactor), L.stream_ptr(),
        )
        L.written(vorticity_field)

    return update_vorticity_from_velocity_forcing_pyst_kernel_3d


def gen_update_vorticity_from_penalised_velocity_pyst_kernel_3d(
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
) -> Callable:
    """``update_vorticity_from_velocity_forcing_3d.py:135-291``:
    ``vorticity[1:-1]^3 += prefactor * curl_h(penalised_velocity - velocity)``."""
    tag = L.dtype_tag(real_t)

    def update_vorticity_from_penalised_velocity_pyst_kernel_3d(
        vorticity_field, penalised_velocity_field, velocity_field, prefactor: float
    ) -> None:
        nz, ny, nx = _shape3(vorticity_field)
        L.call(
            "sopht_update_vorticity_from_penalised_velocity_3d", tag, L.comps(vorticity_field),
            L.strides3(vorticity_field[0]), L.comps(penalised_velocity_field),
            L.strides3(penalised_velocity_field[0]), L.comps(velocity_field),
            L.strides3(velocity_field[0]), nz, ny, nx, float(prefactor), L.stream_ptr(),
        )
        L.written(vorticity_field)

    return update_vorticity_from_penalised_velocity_pyst_kernel_3d


def gen_divergence_pyst_kernel_3d(
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
    reset_ghost_zone: bool = True,
) -> Callable:
    """``divergence_3d.py:13-96``."""
    tag = L.dtype_tag(real_t)

    def divergence_pyst_kernel_3d(divergence, field, inv_dx: float) -> None:
        nz, ny, nx = _shape3(divergence)
        L.call(
            "sopht_divergence_3d", tag, L.ptr(divergence), L.strides3(divergence), L.comps(field),
            L.strides3(field[0]), nz, ny, nx, float(inv_dx), int(bool(reset_ghost_zone)),
            L.stream_ptr(),
        )
        L.written(divergence)

    return divergence_pyst_kernel_3d


# ------------------------------------------------------------------------------ diffusion


def gen_diffusion_flux_pyst_kernel_3d(
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
    field_type: str = "scalar",
    reset_ghost_zone: bool = True,
) -> Callable:
    """``diffusion_flux_3d.py:14-115``."""
    _check_field_type(field_type)
    tag = L.dtype_tag(real_t)
    reset = int(bool(reset_ghost_zone))

    def diffusion_flux_pyst_kernel_3d(diffusion_flux, field, prefactor: float) -> None:
        nz, ny, nx = _shape3(field)
        L.call(
            "sopht_diffusion_flux_3d", tag, L.ptr(diffusion_flux), L.strides3(diffusion_flux),
            L.ptr(field), L.strides3(field), nz, ny, nx, float(prefactor), reset, L.stream_ptr(),
        )
        L.written(diffusion_flux)

    if field_type == "scalar":
        return diffusion_flux_pyst_kernel_3d

    def vector_field_diffusion_flux_pyst_kernel_3d(
        vector_field_diffusion_flux, vector_field, prefactor: float
    ) -> None:
        for i in range(3):
            diffusion_flux_pyst_kernel_3d(vector_field_diffusion_flux[i], vector_field[i], prefactor)

    return vector_field_diffusion_flux_pyst_kernel_3d


def gen_diffusion_timestep_euler_forward_pyst_kernel_3d(
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
    field_type: str = "scalar",
) -> Callable:
    """``diffusion_timestep_3d.py:12-80``: flux with border reset into the scalar buffer,
    then ``field += flux`` (per component for vectors, sharing the buffer)."""
    _check_field_type(field_type)
    tag = L.dtype_tag(real_t)

    def diffusion_timestep_euler_forward_pyst_kernel_3d(field, diffusion_flux, nu_dt_by_dx2: float) -> None:
        nz, ny, nx = _shape3(field)
        L.call(
            "sopht_diffusion_timestep_euler_forward_3d", tag, L.ptr(field), L.strides3(field),
            L.ptr(diffusion_flux), L.strides3(diffusion_flux), nz, ny, nx, float(nu_dt_by_dx2),
            L.stream_ptr(),
        )
        L.written(field, diffusion_flux)

    if field_type == "scalar":
        return diffusion_timestep_euler_forward_pyst_kernel_3d

    def vector_field_diffusion_timestep_euler_forward_pyst_kernel_3d(
        vector_field, diffusion_flux, nu_dt_by_dx2: float
    ) -> None:
        for i in range(3):
            diffusion_timestep_euler_forward_pyst_kernel_3d(vector_field[i], diffusion_flux, nu_dt_by_dx2)

    return vector_field_diffusion_timestep_euler_forward_pyst_kernel_3d


# ------------------------------------------------------------------------------ penalisation


def penalisation_sine_tables(width, dx, x_grid_field, y_grid_field, z_grid_field) -> np.ndarray:
    """Slab factors of ``penalise_field_boundary_3d.py:69-180``, in FP64 from the
    given (host or device) grid fields: per axis x, y, z, ``width`` front factors
    ``sin(pi/(2 w dx) (c_i - c_0))`` then ``width`` back factors ``sin(pi/(2 w dx) (c_end - c_i))``."""

    def host(a):
        return a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)

    xg, yg, zg = host(x_grid_field), host(y_grid_field), host(z_grid_field)
    sp = (np.pi / 2) / (width * float(dx))
    xs = xg[0, 0, :].astype(np.float64)
    ys = yg[0, :, 0].astype(np.float64)
    zs = zg[:, 0, 0].astype(np.float64)
    out = []
    for c in (xs, ys, zs):
        out.extend(np.sin(sp * (c[:width] - c[0])))
        out.extend(np.sin(sp * (c[-1] - c[-width:])))
    return np.asarray(out, dtype=np.float64)


def gen_penalise_field_boundary_pyst_kernel_3d(
    width: int,
    dx: float,
    x_grid_field,
    y_grid_field,
    z_grid_field,
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
    field_type: str = "scalar",
) -> Callable:
    """``penalise_field_boundary_3d.py:13-240``; in place, race-free single pass on the
    device (clamped source times the per-axis sine factors)."""
    if not isinstance(width, int) or width < 0:
        msg = "Invalid width for boundary zone, must be a non-negative integer"
        raise ValueError(msg)
    _check_field_type(field_type)
    tag = L.dtype_tag(real_t)
    if width == 0:
        if field_type == "scalar":

            def penalise_field_boundary_pyst_kernel_3d(field) -> None:
                pass

            return penalise_field_boundary_pyst_kernel_3d

        def penalise_vector_field_boundary_pyst_kernel_3d(vector_field) -> None:
            pass

        return penalise_vector_field_boundary_pyst_kernel_3d

    tables = penalisation_sine_tables(width, dx, x_grid_field, y_grid_field, z_grid_field)
    import ctypes as C

    ctab = (C.c_double * len(tables))(*tables)

    def penalise_field_boundary_pyst_kernel_3d(field) -> None:
        nz, ny, nx = _shape3(field)
        if width == 1:
            # the reference's back-slab source ``field[..., -1:0]`` is empty for width 1, and
            # its broadcast into the one-plane slab raises (penalise_field_boundary_3d.py:192)
            msg = (f"could not broadcast input array from shape ({nz},{ny},0) "
                   f"into shape ({nz},{ny},1)")
            raise ValueError(msg)
        L.call(
            "sopht_penalise_field_boundary_3d", tag, L.ptr(field), L.strides3(field), nz, ny, nx,
            width, ctab, L.stream_ptr(),
        )
        L.written(field)

    if field_type == "scalar":
        return penalise_field_boundary_pyst_kernel_3d

    def penalise_vector_field_boundary_pyst_kernel_3d(vector_field) -> None:
        for i in range(3):
            penalise_field_boundary_pyst_kernel_3d(vector_field[i])

    return penalise_vector_field_boundary_pyst_kernel_3d


# ------------------------------------------------------------------------------ elementwise


def gen_elementwise_sum_pyst_kernel_3d(
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
    field_type: str = "scalar",
) -> Callable:
    """``elementwise_ops_3d.py:13-57``."""
    _check_field_type(field_type)
    tag = L.dtype_tag(real_t)

    def _sum(sum_field, field_1, field_2):
        nz, ny, nx = _shape3(sum_field)
        L.call(
            "sopht_elementwise_sum_3d", tag, L.ptr(sum_field), L.strides3(sum_field), L.ptr(field_1),
            L.strides3(field_1), L.ptr(field_2), L.strides3(field_2), nz, ny, nx, L.stream_ptr(),
        )
        L.written(sum_field)

    def elementwise_sum_pyst_kernel_3d(sum_field, field_1, field_2) -> None:
        if field_type == "vector":
            for i in range(3):
                _sum(sum_field[i], field_1[i], field_2[i])
        else:
            _sum(sum_field, field_1, field_2)

    return elementwise_sum_pyst_kernel_3d


def gen_set_fixed_val_pyst_kernel_3d(
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
    field_type: str = "scalar",
) -> Callable:
    """``elementwise_ops_3d.py:60-119``."""
    _check_field_type(field_type)
    tag = L.dtype_tag(real_t)

    def set_fixed_val_pyst_kernel_3d(field, fixed_val: float) -> None:
        nz, ny, nx = _shape3(field)
        L.call("sopht_set_fixed_val_3d", tag, L.ptr(field), L.strides3(field), nz, ny, nx,
               float(fixed_val), L.stream_ptr())
        L.written(field)

    if field_type == "scalar":
        return set_fixed_val_pyst_kernel_3d

    def vector_field_set_fixed_val_pyst_kernel_3d(vector_field, fixed_vals) -> None:
        for i in range(3):
            set_fixed_val_pyst_kernel_3d(vector_field[i], fixed_vals[i])

    return vector_field_set_fixed_val_pyst_kernel_3d


def gen_elementwise_copy_pyst_kernel_3d(
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
) -> Callable:
    """``elementwise_ops_3d.py:122-143``."""
    tag = L.dtype_tag(real_t)

    def elementwise_copy_pyst_kernel_3d(field, rhs_field) -> None:
        nz, ny, nx = _shape3(field)
        L.call("sopht_elementwise_copy_3d", tag, L.ptr(field), L.strides3(field), L.ptr(rhs_field),
               L.strides3(rhs_field), nz, ny, nx, L.stream_ptr())
        L.written(field)

    return elementwise_copy_pyst_kernel_3d


def gen_elementwise_complex_product_pyst_kernel_3d(
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
) -> Callable:
    """``elementwise_ops_3d.py:146-197`` on complex tensors (interleaved re/im)."""
    tag = L.dtype_tag(real_t)

    def elementwise_complex_product_pyst_kernel_3d(product_field, field_1, field_2) -> None:
        nz, ny, nx = _shape3(product_field)
        L.call("sopht_elementwise_complex_product_3d", tag, L.ptr(product_field),
               L.strides3(product_field), L.ptr(field_1), L.strides3(field_1), L.ptr(field_2),
               L.strides3(field_2), nz, ny, nx, L.stream_ptr())
        L.written(product_field)

    return elementwise_complex_product_pyst_kernel_3d


def gen_set_fixed_val_at_boundaries_pyst_kernel_3d(
    real_t: type,
    width: int,
    num_threads: bool | int = False,
    field_type: str = "scalar",
) -> Callable:
    """``elementwise_ops_3d.py:200-268``."""
    if not isinstance(width, int) or width <= 0:
        msg = "Invalid width for boundary zone, must be a positive integer"
        raise ValueError(msg)
    _check_field_type(field_type)
    tag = L.dtype_tag(real_t)

    def set_fixed_val_at_boundaries_pyst_kernel_3d(field, fixed_val: float) -> None:
        nz, ny, nx = _shape3(field)
        L.call("sopht_set_fixed_val_at_boundaries_3d", tag, L.ptr(field), L.strides3(field), nz, ny,
               nx, width, float(fixed_val), L.stream_ptr())
        L.written(field)

    if field_type == "scalar":
        return set_fixed_val_at_boundaries_pyst_kernel_3d

    def vector_field_set_fixed_val_at_boundaries_pyst_kernel_3d(vector_field, fixed_vals) -> None:
        for i in range(3):
            set_fixed_val_at_boundaries_pyst_kernel_3d(vector_field[i], fixed_vals[i])

    return vector_field_set_fixed_val_at_boundaries_pyst_kernel_3d


def gen_add_fixed_val_pyst_kernel_3d(
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
    field_type: str = "scalar",
) -> Callable:
    """``elementwise_ops_3d.py:271-334``."""
    _check_field_type(field_type)
    tag = L.dtype_tag(real_t)

    def add_fixed_val_pyst_kernel_3d(sum_field, field, fixed_val: float) -> None:
        nz, ny, nx = _shape3(field)
        L.call("sopht_add_fixed_val_3d", tag, L.ptr(sum_field), L.strides3(sum_field), L.ptr(field),
               L.strides3(field), nz, ny, nx, float(fixed_val), L.stream_ptr())
        L.written(sum_field)

    if field_type == "scalar":
        return add_fixed_val_pyst_kernel_3d

    def vector_field_add_fixed_val_pyst_kernel_3d(sum_field, vector_field, fixed_vals) -> None:
        for i in range(3):
            add_fixed_val_pyst_kernel_3d(sum_field[i], vector_field[i], float(fixed_vals[i]))

    return vector_field_add_fixed_val_pyst_kernel_3d


def gen_elementwise_saxpby_pyst_kernel_3d(
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
    field_type: str = "scalar",
) -> Callable:
    """``elementwise_ops_3d.py:337-387``: ``sum = a * field_1 + b * field_2``."""
    _check_field_type(field_type)
    tag = L.dtype_tag(real_t)

    def _saxpby(s, f1, f2, a, b):
        nz, ny, nx = _shape3(s)
        L.call("sopht_elementwise_saxpby_3d", tag, L.ptr(s), L.strides3(s), L.ptr(f1), L.strides3(f1),
               L.ptr(f2), L.strides3(f2), nz, ny, nx, float(a), float(b), L.stream_ptr())
        L.written(s)

    def elementwise_saxpby_pyst_kernel_3d(sum_field, field_1, field_2, field_1_prefac, field_2_prefac) -> None:
        if field_type == "vector":
            for i in range(3):
                _saxpby(sum_field[i], field_1[i], field_2[i], field_1_prefac, field_2_prefac)
        else:
            _saxpby(sum_field, field_1, field_2, field_1_prefac, field_2_prefac)

    return elementwise_saxpby_pyst_kernel_3d


def gen_elementwise_cross_product_pyst_kernel_3d(
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
) -> Callable:
    """``elementwise_ops_3d.py:390-449``."""
    tag = L.dtype_tag(real_t)

    def elementwise_cross_product_pyst_kernel_3d(result_field, field_1, field_2) -> None:
        nz, ny, nx = _shape3(result_field)
        L.call("sopht_elementwise_cross_product_3d", tag, L.comps(result_field),
               L.strides3(result_field[0]), L.comps(field_1), L.strides3(field_1[0]),
               L.comps(field_2), L.strides3(field_2[0]), nz, ny, nx, L.stream_ptr())
        L.written(result_field)

    return elementwise_cross_product_pyst_kernel_3d


# ------------------------------------------------------------------------------ off-step (ENO3, stretching)


def gen_advection_flux_conservative_eno3_pyst_kernel_3d(
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
) -> Callable:
    """``advection_flux_3d.py:12-233``: accumulates into ``advection_flux`` on ``[2:-2]^3``."""
    tag = L.dtype_tag(real_t)

    def advection_flux_conservative_eno3_pyst_kernel_3d(advection_flux, field, velocity, inv_dx: float) -> None:
        nz, ny, nx = _shape3(field)
        L.call("sopht_advection_flux_conservative_eno3_3d", tag, L.ptr(advection_flux),
               L.strides3(advection_flux), L.ptr(field), L.strides3(field), L.comps(velocity),
               L.strides3(velocity[0]), nz, ny, nx, float(inv_dx), L.stream_ptr())
        L.written(advection_flux)

    return advection_flux_conservative_eno3_pyst_kernel_3d


def gen_advection_flux_conservative_eno3_adjoint_kernel_3d(real_t: type) -> Callable:
    """Transpose of the ENO3 flux above (``advection_flux_3d.py:12-233``; the reference has no
    adjoint), with each face's upwind stencil choice frozen at ``velocity``.

    ``kernel(field_adj, velocity_adj, flux_adj, field, velocity, inv_dx)`` accumulates
    ``J_field^T flux_adj`` into ``field_adj`` and ``J_velocity^T flux_adj`` into
    ``velocity_adj`` (either may be None). ``flux_adj`` is read on ``[2:-2]^3``, the cells
    the forward writes."""
    tag = L.dtype_tag(real_t)

    def advection_flux_conservative_eno3_adjoint_kernel_3d(
        field_adj, velocity_adj, flux_adj, field, velocity, inv_dx: float
    ) -> None:
        nz, ny, nx = _shape3(field)
        L.call("sopht_advection_flux_conservative_eno3_adjoint_3d", tag, L.ptr(field_adj),
               L.strides3(field_adj) if field_adj is not None else None,
               L.comps(velocity_adj) if velocity_adj is not None else None,
               L.strides3(velocity_adj[0]) if velocity_adj is not None else None,
               L.ptr(flux_adj), L.strides3(flux_adj), L.ptr(field), L.strides3(field),
               L.comps(velocity), L.strides3(velocity[0]), nz, ny, nx, float(inv_dx),
               L.stream_ptr())
        written = [t for t in (field_adj, velocity_adj) if t is not None]
        L.written(*written)

    return advection_flux_conservative_eno3_adjoint_kernel_3d


def eno3_advection_flux(field, velocity, inv_dx: float):
    """Differentiable ENO3 conservative advection flux (a ``torch.autograd.Function``).

    Returns a new flux field: zero outside ``[2:-2]^3`` and the reference's ENO3 flux inside.
    The backward runs the transpose kernel with the forward's stencil choices (they depend
    on ``velocity`` only), so gradients flow to both ``field`` and ``velocity``; the switch
    itself is piecewise constant and contributes nothing."""
    import torch

    real_t = np.float32 if field.dtype == torch.float32 else np.float64
    forward_kernel = gen_advection_flux_conservative_eno3_pyst_kernel_3d(real_t)
    adjoint_kernel = gen_advection_flux_conservative_eno3_adjoint_kernel_3d(real_t)

    class _ENO3Flux(torch.autograd.Function):
        @staticmethod
        def forward(ctx, f, u):
            f, u = f.contiguous(), u.contiguous()
            flux = torch.zeros_like(f)
            forward_kernel(flux, f, u, inv_dx)
            ctx.save_for_backward(f, u)
            return flux

        @staticmethod
        def backward(ctx, grad_flux):
            f, u = ctx.saved_tensors
            grad_flux = grad_flux.contiguous()
            gf = torch.zeros_like(f) if ctx.needs_input_grad[0] else None
            gu = torch.zeros_like(u) if ctx.needs_input_grad[1] else None
            if gf is not None or gu is not None:
                adjoint_kernel(gf, gu, grad_flux, f, u, inv_dx)
            return gf, gu

    return _ENO3Flux.apply(field, velocity)


def gen_advection_timestep_euler_forward_conservative_eno3_pyst_kernel_3d(
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
    field_type: str = "scalar",
) -> Callable:
    """``advection_timestep_3d.py:12-99``: zero flux, ENO3 flux with ``inv_dx=-dt_by_dx``,
    ``field += flux`` (per component for vectors)."""
    _check_field_type(field_type)
    set_val = gen_set_fixed_val_pyst_kernel_3d(real_t)
    flux_kernel = gen_advection_flux_conservative_eno3_pyst_kernel_3d(real_t)
    add = gen_elementwise_sum_pyst_kernel_3d(real_t)

    def advection_timestep_euler_forward_conservative_eno3_pyst_kernel_3d(
        field, advection_flux, velocity, dt_by_dx: float
    ) -> None:
        set_val(field=advection_flux, fixed_val=0)
        flux_kernel(advection_flux=advection_flux, field=field, velocity=velocity, inv_dx=-dt_by_dx)
        add(sum_field=field, field_1=field, field_2=advection_flux)

    if field_type == "scalar":
        return advection_timestep_euler_forward_conservative_eno3_pyst_kernel_3d

    def vector_field_advection_timestep_euler_forward_conservative_eno3_pyst_kernel_3d(
        vector_field, advection_flux, velocity, dt_by_dx: float
    ) -> None:
        for i in range(3):
            advection_timestep_euler_forward_conservative_eno3_pyst_kernel_3d(
                vector_field[i], advection_flux, velocity, dt_by_dx
            )

    return vector_field_advection_timestep_euler_forward_conservative_eno3_pyst_kernel_3d


def gen_vorticity_stretching_flux_pyst_kernel_3d(
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
) -> Callable:
    """``vorticity_stretching_flux_3d.py:13-107``."""
    tag = L.dtype_tag(real_t)

    def vorticity_stretching_flux_pyst_kernel_3d(
        vorticity_stretching_flux_field, vorticity_field, velocity_field, prefactor: float
    ) -> None:
        nz, ny, nx = _shape3(vorticity_field)
        L.call("sopht_vorticity_stretching_flux_3d", tag, L.comps(vorticity_stretching_flux_field),
               L.strides3(vorticity_stretching_flux_field[0]), L.comps(vorticity_field),
               L.strides3(vorticity_field[0]), L.comps(velocity_field), L.strides3(velocity_field[0]),
               nz, ny, nx, float(prefactor), L.stream_ptr())
        L.written(vorticity_stretching_flux_field)

    return vorticity_stretching_flux_pyst_kernel_3d


def gen_vorticity_stretching_timestep_euler_forward_pyst_kernel_3d(
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
) -> Callable:
    """``vorticity_stretching_timestep_3d.py:10-55``."""
    flux = gen_vorticity_stretching_flux_pyst_kernel_3d(real_t)
    add = gen_elementwise_sum_pyst_kernel_3d(real_t, field_type="vector")

    def vorticity_stretching_timestep_euler_forward_pyst_kernel_3d(
        vorticity_field, velocity_field, vorticity_stretching_flux_field, dt_by_2_dx: float
    ) -> None:
        flux(vorticity_stretching_flux_field, vorticity_field, velocity_field, dt_by_2_dx)
        add(vorticity_field, vorticity_field, vorticity_stretching_flux_field)

    return vorticity_stretching_timestep_euler_forward_pyst_kernel_3d


def gen_vorticity_stretching_timestep_ssprk3_pyst_kernel_3d(
    real_t: type,
    midstep_buffer_vector_field,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
) -> Callable:
    """``vorticity_stretching_timestep_3d.py:58-158`` (both mid-step views alias
    ``midstep_buffer_vector_field``; third stage prefactor ``dt_by_2_dx * 0.5``)."""
    flux = gen_vorticity_stretching_flux_pyst_kernel_3d(real_t)
    add = gen_elementwise_sum_pyst_kernel_3d(real_t, field_type="vector")
    saxpby = gen_elementwise_saxpby_pyst_kernel_3d(real_t, field_type="vector")
    m = midstep_buffer_vector_field

    def vorticity_stretching_timestep_ssprk3_pyst_kernel_3d(
        vorticity_field, velocity_field, vorticity_stretching_flux_field, dt_by_2_dx: float
    ) -> None:
        f = vorticity_stretching_flux_field
        flux(f, vorticity_field, velocity_field, dt_by_2_dx)
        add(m, vorticity_field, f)
        flux(f, m, velocity_field, dt_by_2_dx)
        add(m, m, f)
        saxpby(m, vorticity_field, m, 0.75, 0.25)
        flux(f, m, velocity_field, dt_by_2_dx * 0.5)
        add(m, m, f)
        saxpby(vorticity_field, vorticity_field, m, 1.0 / 3.0, 2.0 / 3.0)

    return vorticity_stretching_timestep_ssprk3_pyst_kernel_3d


def gen_laplacian_filter_kernel_3d(
    filter_order: int,
    filter_flux_buffer,
    field_buffer,
    real_t: type,
    num_threads: bool | int = False,
    fixed_grid_size: tuple[int, int, int] | bool = False,
    field_type: str = "scalar",
    filter_type: str = "multiplicative",
    filter_flux_buffer_boundary_width: int = 1,
) -> Callable:
    """``laplacian_filter_3d.py:13-194``: 1D Laplacian filters along x, y, z (Jeanmart and
    Winckelmans 2007), ``"multiplicative"`` or ``"convolution"``, applied in place. The closure
    keeps the caller's two scalar buffers (so it is not re-entrant across users of the same
    buffers, as in the reference)."""
    if not isinstance(filter_order, int) or filter_order < 0:
        msg = "Invalid filter order, must be a non-negative integer"
        raise ValueError(msg)
    if not isinstance(filter_flux_buffer_boundary_width, int) or filter_flux_buffer_boundary_width <= 0:
        msg = "Invalid value for filter flux buffer boundary zone, must be a positive integer"
        raise ValueError(msg)
    types = {"multiplicative": 0, "convolution": 1}
    if filter_type not in types:
        msg = "Invalid filter type"
        raise ValueError(msg)
    if field_type not in ("scalar", "vector"):
        msg = "Invalid field type"
        raise ValueError(msg)
    tag = L.dtype_tag(real_t)
    ftype = types[filter_type]
    width = int(filter_flux_buffer_boundary_width)

    def scalar_field_filter_kernel_3d(scalar_field) -> None:
        nz, ny, nx = _shape3(scalar_field)
        L.call(
            "sopht_laplacian_filter_3d", tag, int(filter_order), ftype, L.ptr(scalar_field),
            L.strides3(scalar_field), nz, ny, nx, L.ptr(filter_flux_buffer), L.strides3(filter_flux_buffer),
            L.ptr(field_buffer), L.strides3(field_buffer), width, L.stream_ptr(),
        )
        L.written(scalar_field, filter_flux_buffer, field_buffer)

    if field_type == "scalar":
        return scalar_field_filter_kernel_3d

    def vector_field_filter_kernel_3d(vector_field) -> None:
        for i in range(3):
            scalar_field_filter_kernel_3d(scalar_field=vector_field[i])

    return vector_field_filter_kernel_3d
