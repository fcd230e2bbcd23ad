"""Drop-in ``PassiveTransportFlowSimulator`` (3D), ``passive_transport_flow_simulators.py:14-136``.

Advection-diffusion of a passive scalar or vector ``primary_field`` by a given
``velocity_field``. The reference runs two factory kernels per component (the conservative
ENO3 advection step, then the Euler diffusion step, sharing ``buffer_scalar_field`` as
flux buffer); here one ``time_step`` is one call of ``sopht_passive_transport_step_3d``:
a fused sweep that advects and diffuses all components together, reading the velocity
once, with the advected field kept on chip. ``compute_stable_timestep`` reduces
``sum_i |u_i|`` on the device (``sopht_velocity_abs_sum_max_3d``, which also leaves the
per-cell sums in ``buffer_scalar_field`` as the reference does) and reads back one scalar.

Public attributes, constructor arguments and error messages are the reference's; the
fields are torch tensors on the HIP device.
"""

from __future__ import annotations

import numpy as np

from .. import _lib as L
from .flow import FlowSimulator

_PARTIALS = 4096  # partial maxima of the CFL reduction (one per workgroup, at most)


class PassiveTransportFlowSimulator(FlowSimulator):
    """Passive transport (advection-diffusion) of ``primary_field`` by ``velocity_field``."""

    def __init__(
        self,
        kinematic_viscosity,
        grid_dim,
        grid_size,
        x_range,
        cfl=0.1,
        real_t=np.float32,
        num_threads=1,
        time=0.0,
        field_type="scalar",
    ):
        # argument checks in the reference's order and wording (:45-54)
        if field_type not in ("scalar", "vector"):
            msg = "Invalid field type. Supported values include 'scalar' and 'vector'"
            raise ValueError(msg)
        if grid_dim == 2 and field_type == "vector":
            msg = "Passive transport of vector 2D fields not supported yet."
            raise ValueError(msg)
        self.kinematic_viscosity = kinematic_viscosity
        self.cfl = cfl
        self.field_type = field_type
        super().__init__(grid_dim, grid_size, x_range, real_t, num_threads, time)

    def _init_fields(self):
        import torch

        dev = torch.device("cuda", torch.cuda.current_device())
        make = lambda *shape: torch.zeros(shape, dtype=L.torch_dtype(self.real_t), device=dev)  # noqa: E731
        self._ncomp = 1 if self.field_type == "scalar" else self.grid_dim
        comp_shape = tuple(self.grid_size)
        self.primary_field = make(*comp_shape) if self._ncomp == 1 else make(self._ncomp, *comp_shape)
        self.velocity_field = make(self.grid_dim, *comp_shape)
        self.buffer_scalar_field = make(*comp_shape)
        # the sweep writes the new field out of place (neighbouring tiles still read the old
        # one), then copies it back: one scratch field per component
        self._next = make(self._ncomp, *comp_shape)
        self._partials = make(_PARTIALS)
        self._vmax = make(1)

    def _compile_kernels(self):
        # nothing to generate: the step is one library entry point, bound here
        self._tag = L.dtype_tag(self.real_t)
        L.lib()

    def _finalise_flow_time_step(self):
        self._flow_time_step = self._transport

    def _component_ptrs(self, t, count):
        ptrs = [t.data_ptr()] if count == 1 and t.dim() == 3 else [t[c].data_ptr() for c in range(count)]
        return L._p3(*(ptrs + [0] * (3 - len(ptrs))))

    def _transport(self, dt, **kwargs):
        """One advection-diffusion step of ``dt`` (``:100-113``); extra keywords are ignored
        as in the reference."""
        for name, t in (("primary_field", self.primary_field), ("velocity_field", self.velocity_field)):
            if not t.is_contiguous():
                msg = f"{name} must be contiguous"
                raise ValueError(msg)
        nz, ny, nx = self.grid_size
        # real_t(dt / dx) and real_t(nu dt / dx / dx), as the reference forms them
        dt_by_dx = self.real_t(dt / self.dx)
        nu_dt_by_dx2 = self.real_t(self.kinematic_viscosity * dt / self.dx / self.dx)
        L.call(
            "sopht_passive_transport_step_3d", self._tag, self._ncomp,
            self._component_ptrs(self.primary_field, self._ncomp), self._component_ptrs(self._next, self._ncomp),
            L.comps(self.velocity_field), nz, ny, nx, float(dt_by_dx), float(nu_dt_by_dx2), L.stream_ptr(),
        )
        L.written(self.primary_field, self._next)

    def compute_stable_timestep(self, dt_prefac=1.0):
        """``:120-136``: the advection-diffusion limit on ``dt``, times ``dt_prefac``."""
        return dt_prefac * compute_advection_diffusion_stable_timestep(
            self.velocity_field, self.buffer_scalar_field, self.grid_dim, self.dx, self.cfl,
            self.kinematic_viscosity, self.real_t, _scratch=(self._partials, self._vmax),
        )


def compute_advection_diffusion_stable_timestep(
    velocity_field, velocity_magnitude_field, grid_dim, dx, cfl, kinematic_viscosity, real_t=np.float32,
    _scratch=None,
):
    """``passive_transport_flow_simulators.py:139-155`` on device tensors: the CFL limit
    ``cfl dx / (max sum_i |u_i| + tol)`` against the diffusion limit
    ``0.9 dx^2 / (2 d) / nu + tol`` with ``tol = 10 eps(real_t)``. The per-cell sums land in
    ``velocity_magnitude_field``; the maximum is reduced on the device (one scalar read back)."""
    import torch

    v = velocity_field
    if not v.is_contiguous() or not velocity_magnitude_field.is_contiguous():
        msg = "velocity_field and velocity_magnitude_field must be contiguous"
        raise ValueError(msg)
    if _scratch is None:
        _scratch = (torch.empty(_PARTIALS, dtype=v.dtype, device=v.device),
                    torch.empty(1, dtype=v.dtype, device=v.device))
    partials, vmax = _scratch
    n = velocity_magnitude_field.numel()
    L.call("sopht_velocity_abs_sum_max_3d", L.dtype_tag(real_t), L.comps(v), n, L.ptr(velocity_magnitude_field),
           L.ptr(partials), int(partials.numel()), L.ptr(vmax), L.stream_ptr())
    L.written(velocity_magnitude_field)
    tol = 10 * np.finfo(real_t).eps
    advective = cfl * dx / (real_t(vmax.item()) + tol)
    diffusive = 0.9 * dx**2 / (2 * grid_dim) / kinematic_viscosity + tol
    return min(advective, diffusive)
