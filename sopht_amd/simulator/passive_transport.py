"""Drop-in ``PassiveTransportFlowSimulator`` (3D): advection-diffusion of a passive scalar or
vector field, ``passive_transport_flow_simulators.py:14-136``.

One ``time_step`` is the reference's sequence on device tensors: the conservative ENO3
advection step (``advection_timestep_3d.py:12-99``) and the Euler diffusion step
(``diffusion_timestep_3d.py:12-80``), both through the drop-in stencil factories, sharing
``buffer_scalar_field`` as their flux buffer.
"""

from __future__ import annotations

import numpy as np

from .. import _lib as L
from .flow import FlowSimulator, compute_advection_diffusion_stable_timestep


class PassiveTransportFlowSimulator(FlowSimulator):
    """Advection-diffusion of ``primary_field`` by ``velocity_field``."""

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
        if field_type not in ["scalar", "vector"]:
            msg = "Invalid field type. Supported values include 'scalar' and 'vector'"
            raise ValueError(msg)
        if grid_dim == 2 and field_type == "vector":
            msg = "Passive transport of vector 2D fields not supported yet."
            raise ValueError(msg)
        self.kinematic_viscosity = kinematic_viscosity
        self.cfl = cfl
        self.field_type = field_type
        super().__init__(grid_dim, grid_size, x_range, real_t, num_threads, time)

    # :57-67
    def _init_fields(self):
        import torch

        dev = torch.device("cuda", torch.cuda.current_device())
        tdt = L.torch_dtype(self.real_t)
        shape = self.grid_size if self.field_type == "scalar" else (self.grid_dim, *self.grid_size)
        self.primary_field = torch.zeros(shape, dtype=tdt, device=dev)
        self.velocity_field = torch.zeros((self.grid_dim, *self.grid_size), dtype=tdt, device=dev)
        # the same buffer for the advection and diffusion fluxes
        self.buffer_scalar_field = torch.zeros(self.grid_size, dtype=tdt, device=dev)

    # :69-98
    def _compile_kernels(self):
        from ..numeric import eulerian_grid_ops as spne

        self._diffusion_timestep = spne.gen_diffusion_timestep_euler_forward_pyst_kernel_3d(
            real_t=self.real_t, fixed_grid_size=self.grid_size, num_threads=self.num_threads,
            field_type=self.field_type,
        )
        self._advection_timestep = spne.gen_advection_timestep_euler_forward_conservative_eno3_pyst_kernel_3d(
            real_t=self.real_t, fixed_grid_size=self.grid_size, num_threads=self.num_threads,
            field_type=self.field_type,
        )

    # :100-113
    def _advection_and_diffusion_time_step(self, dt, **kwargs):
        del kwargs
        self._advection_timestep(
            self.primary_field,
            advection_flux=self.buffer_scalar_field,
            velocity=self.velocity_field,
            dt_by_dx=self.real_t(dt / self.dx),
        )
        self._diffusion_timestep(
            self.primary_field,
            diffusion_flux=self.buffer_scalar_field,
            nu_dt_by_dx2=self.real_t(self.kinematic_viscosity * dt / self.dx / self.dx),
        )

    # :115-118
    def _finalise_flow_time_step(self):
        self._flow_time_step = self._advection_and_diffusion_time_step

    # :120-136
    def compute_stable_timestep(self, dt_prefac=1.0):
        dt = compute_advection_diffusion_stable_timestep(
            velocity_field=self.velocity_field,
            velocity_magnitude_field=self.buffer_scalar_field,
            grid_dim=self.grid_dim,
            dx=self.dx,
            cfl=self.cfl,
            kinematic_viscosity=self.kinematic_viscosity,
            real_t=self.real_t,
        )
        return dt * dt_prefac
