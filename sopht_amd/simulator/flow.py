"""Drop-in ``UnboundedNavierStokesFlowSimulator3D`` whose step runs as fused HIP sweeps.

Mirrors ``sopht/simulator/flow/navier_stokes_flow_simulators.py:226-522`` (3D class),
``flow_simulators.py:9-133`` (base) and ``flow_simulators_3d.py:8-65`` (factory). Field
attributes (``vorticity_field``, ``velocity_field``, ``buffer_vector_field``,
``buffer_scalar_field``, ``stream_func_field``, ``eul_grid_forcing_field``) are
``(3, Nz, Ny, Nx)`` torch tensors on the HIP device, allocated once; ``position_field``
stays a host numpy array as in the reference. One call of ``time_step`` is one
``sopht_ns3d_step`` (three fused stencil sweeps, the pruned Poisson passes, the velocity
sweep) on torch's current stream; nothing in it synchronises.
"""

from __future__ import annotations

import ctypes as C
import logging

import numpy as np

from .. import _lib as L

logger = logging.getLogger(__name__)

_zeros_3 = np.zeros(3, dtype=np.float64)


class FlowSimulator:
    """Base flow simulator (``flow_simulators.py:9-133``)."""

    def __init__(self, grid_dim, grid_size, x_range, real_t=np.float32, num_threads=1, time=0.0):
        if grid_dim not in [2, 3]:
            msg = "Invalid grid dimensions. Supported values include 2 and 3."
            raise ValueError(msg)
        if grid_dim == 2:
            msg = "2D flow is outside this backend (3D flow step only)"
            raise NotImplementedError(msg)
        self.grid_dim = grid_dim
        self.grid_size = tuple(int(g) for g in grid_size)
        self.x_range = x_range
        self.real_t = real_t
        self.num_threads = num_threads
        self.time = time
        self._init_domain()
        self._init_fields()
        self._compile_kernels()
        self._finalise_flow_time_step()

    def _init_domain(self):
        """``flow_simulators.py:50-95``."""
        grid_size_z, grid_size_y, grid_size_x = self.grid_size
        self.dx = self.real_t(self.x_range / grid_size_x)
        eul_grid_shift = self.dx / 2.0
        x = np.linspace(eul_grid_shift, self.x_range - eul_grid_shift, grid_size_x).astype(self.real_t)
        self.y_range = self.x_range * grid_size_y / grid_size_x
        self.z_range = self.x_range * grid_size_z / grid_size_x
        y = np.linspace(eul_grid_shift, self.y_range - eul_grid_shift, grid_size_y).astype(self.real_t)
        z = np.linspace(eul_grid_shift, self.z_range - eul_grid_shift, grid_size_z).astype(self.real_t)
        self.position_field = np.flipud(np.array(np.meshgrid(z, y, x, indexing="ij")))
        logger.info(
            "\n3D flow domain initialized with:\nX axis from 0.0 to %s\nY axis from 0.0 to %s"
            "\nZ axis from 0.0 to %s",
            self.x_range,
            self.y_range,
            self.z_range,
        )

    def _update_simulator_time(self, dt):
        self.time += dt

    def time_step(self, dt, **kwargs):
        """``flow_simulators.py:130-133``."""
        self._flow_time_step(dt=dt, **kwargs)
        self._update_simulator_time(dt=dt)


class UnboundedNavierStokesFlowSimulator3D(FlowSimulator):
    """3D unbounded Navier-Stokes flow simulator (greens-function Poisson solver)."""

    def __init__(
        self,
        grid_size,
        x_range,
        kinematic_viscosity,
        cfl=0.1,
        real_t=np.float32,
        num_threads=1,
        time=0.0,
        with_forcing=False,
        with_free_stream_flow=False,
        filter_vorticity=False,
        flow_density=1.0,
        poisson_solver_type="greens_function_convolution",
        **kwargs,
    ):
        self.kinematic_viscosity = kinematic_viscosity
        self.cfl = cfl
        self.with_forcing = with_forcing
        self.with_free_stream_flow = with_free_stream_flow
        self.penalty_zone_width = kwargs.get("penalty_zone_width", 2)
        self.filter_vorticity = filter_vorticity
        self.filter_setting_dict = None
        if filter_vorticity:
            # :274-292: default order 2, multiplicative; custom settings via filter_setting_dict
            default_filter_setting_dict = {"order": 2, "type": "multiplicative"}
            self.filter_setting_dict = kwargs.get("filter_setting_dict", default_filter_setting_dict)
            logger.info(
                "Vorticity filtering is turned on: order = %s, type = %s",
                self.filter_setting_dict["order"], self.filter_setting_dict["type"],
            )
        self.flow_density = flow_density
        supported = ["greens_function_convolution", "fast_diagonalisation"]
        self.poisson_solver_type = poisson_solver_type
        if poisson_solver_type not in supported:
            msg = "Invalid Poisson solver type given"
            raise ValueError(msg)
        super().__init__(
            grid_dim=3, grid_size=grid_size, x_range=x_range, real_t=real_t,
            num_threads=num_threads, time=time,
        )

    # :314-325
    def _init_fields(self):
        import torch

        self._device = torch.device("cuda", torch.cuda.current_device())
        tdt = L.torch_dtype(self.real_t)
        shape = (self.grid_dim, *self.grid_size)
        self.vorticity_field = torch.zeros(shape, dtype=tdt, device=self._device)
        self.velocity_field = torch.zeros_like(self.vorticity_field)
        self.buffer_vector_field = torch.zeros_like(self.vorticity_field)
        self.buffer_scalar_field = self.buffer_vector_field[0]
        self.stream_func_field = torch.zeros_like(self.vorticity_field)
        if self.with_forcing:
            self.eul_grid_forcing_field = torch.zeros_like(self.velocity_field)
        self._max_out = torch.zeros((), dtype=tdt, device=self._device)

    # :328-440
    def _compile_kernels(self):
        from ..numeric import eulerian_grid_ops as spne

        nz, ny, nx = self.grid_size
        plan = C.c_void_p()
        # :337-358: the free-space solver, or the Neumann fast-diagonalisation one
        solver_type = 1 if self.poisson_solver_type == "fast_diagonalisation" else 0
        L.call(
            "sopht_ns3d_create_with_poisson_solver", L.dtype_tag(self.real_t), nz, ny, nx,
            float(self.x_range), float(self.kinematic_viscosity), float(self.flow_density),
            int(self.penalty_zone_width), int(bool(self.with_forcing)),
            int(bool(self.with_free_stream_flow)), solver_type, L.stream_ptr(), C.byref(plan),
        )
        self._plan = plan
        self._lib = L.lib()
        # :396-408 filter applied in the step after diffusion (buffer components 0 and 1)
        if self.filter_vorticity and self.filter_setting_dict is not None:
            order = self.filter_setting_dict["order"]
            ftype = self.filter_setting_dict["type"]
            if not isinstance(order, int) or order < 0:
                msg = "Invalid filter order, must be a non-negative integer"
                raise ValueError(msg)
            if ftype not in ("multiplicative", "convolution"):
                msg = "Invalid filter type"
                raise ValueError(msg)
            L.call("sopht_ns3d_set_filter", plan, int(order), 0 if ftype == "multiplicative" else 1)
        # standalone drop-ins used outside the fused step
        self._compute_divergence = spne.gen_divergence_pyst_kernel_3d(real_t=self.real_t)

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            self._lib.sopht_ns3d_destroy(plan)
            self._plan = None

    # :443-447
    def _finalise_flow_time_step(self):
        self._flow_time_step = self._navier_stokes_time_step
        if self.with_forcing:
            self._flow_time_step = self._navier_stokes_with_forcing_time_step

    def _fs(self, free_stream_velocity):
        fs = np.asarray(free_stream_velocity, dtype=np.float64).reshape(3)
        return (C.c_double * 3)(*fs)

    def _check_penalty_width(self):
        if self.penalty_zone_width == 1:
            # as the reference's penalisation kernel, whose back-slab source is an empty
            # slice for width 1 (penalise_field_boundary_3d.py:192)
            nz, ny, nx = self.grid_size
            msg = f"could not broadcast input array from shape ({nz},{ny},0) into shape ({nz},{ny},1)"
            raise ValueError(msg)

    def _step(self, dt, free_stream_velocity, forcing):
        self._check_penalty_width()
        L.call(
            "sopht_ns3d_step", self._plan, L.ptr(self.vorticity_field), L.ptr(self.velocity_field),
            L.ptr(self.stream_func_field), L.ptr(forcing), L.ptr(self.buffer_vector_field),
            float(dt), self._fs(free_stream_velocity), L.stream_ptr(),
        )
        L.written(self.vorticity_field, self.velocity_field, self.stream_func_field, forcing)
        self._velocity_version = self.velocity_field._version

    # :449-485
    def _navier_stokes_time_step(self, dt, free_stream_velocity=_zeros_3):
        self._step(dt, free_stream_velocity, None)

    # :487-498
    def _navier_stokes_with_forcing_time_step(self, dt, free_stream_velocity=_zeros_3):
        self._step(dt, free_stream_velocity, self.eul_grid_forcing_field)

    def trajectory_step(self, vorticity_in, velocity_in, forcing, vorticity_out, velocity_out, dt,
                        free_stream_velocity=_zeros_3, vorticity1_out=None):
        """The step of ``time_step`` out of place, for recording trajectories (no counterpart in
        the reference): reads the state (vorticity_in, velocity_in, forcing), writes the next
        state into (vorticity_out, velocity_out) and leaves forcing unchanged; optionally
        ``vorticity1_out`` receives omega1 = omega + dt/(2 dx rho) curl_h(F). The simulator's
        own fields and time are untouched except ``stream_func_field`` (scratch)."""
        L.call(
            "sopht_ns3d_trajectory_step", self._plan, L.ptr(vorticity_in), L.ptr(velocity_in),
            L.ptr(forcing), L.ptr(vorticity_out), L.ptr(velocity_out), L.ptr(self.stream_func_field),
            L.ptr(vorticity1_out), float(dt), self._fs(free_stream_velocity), L.stream_ptr(),
        )
        L.written(vorticity_out, velocity_out, self.stream_func_field, vorticity1_out)

    def recover_velocity(self, free_stream_velocity=_zeros_3):
        """Velocity recovery alone (``:475-485``): penalise vorticity, Poisson, curl, free stream."""
        self._check_penalty_width()
        L.call(
            "sopht_ns3d_recover_velocity", self._plan, L.ptr(self.vorticity_field),
            L.ptr(self.velocity_field), L.ptr(self.stream_func_field), self._fs(free_stream_velocity),
            L.stream_ptr(),
        )
        L.written(self.vorticity_field, self.velocity_field, self.stream_func_field)
        self._velocity_version = self.velocity_field._version

    # :501-512
    def compute_stable_timestep(self, dt_prefac=1.0):
        return dt_prefac * self._stable_timestep_from_max(self.max_abs_velocity_sum())

    def max_abs_velocity_sum(self):
        """``max over cells of sum_i |u_i|`` (one readback). When the velocity field is the one
        the last step wrote (torch's version counter of the tensor is unchanged), the maxima
        its velocity sweep kept per workgroup are reduced instead of re-reading the field.
        Every entry point of this package that writes a tensor through raw pointers (the
        drop-in kernels, the solvers, the batched step) moves that counter; code that writes
        the velocity field through its own raw pointers calls ``sopht_amd._lib.written`` on it."""
        lib = self._lib
        rc = -1
        if getattr(self, "_velocity_version", None) == self.velocity_field._version:
            rc = lib.sopht_ns3d_last_velocity_max(self._plan, L.ptr(self.velocity_field),
                                                  L.ptr(self._max_out), L.stream_ptr())
        if rc != 0:
            L.call("sopht_ns3d_max_abs_velocity_sum", self._plan, L.ptr(self.velocity_field),
                   L.ptr(self._max_out), L.stream_ptr())
        return self.real_t(self._max_out.item())

    def _stable_timestep_from_max(self, vmax):
        """``passive_transport_flow_simulators.py:139-155`` given the velocity maximum."""
        real_t = self.real_t
        tol = 10 * np.finfo(real_t).eps
        return min(
            self.cfl * self.dx / (vmax + tol),
            0.9 * self.dx**2 / (2 * self.grid_dim) / self.kinematic_viscosity + tol,
        )

    # :514-522
    def get_vorticity_divergence_l2_norm(self):
        import torch

        divergence_field = self.buffer_scalar_field
        self._compute_divergence(divergence=divergence_field, field=self.vorticity_field, inv_dx=(1.0 / self.dx))
        return float(torch.linalg.vector_norm(divergence_field.double()).item()) * self.dx ** (self.grid_dim / 2.0)


def create_unbounded_flow_simulator_3d(
    grid_size,
    x_range,
    kinematic_viscosity,
    cfl=0.1,
    flow_type="navier_stokes",
    real_t=np.float32,
    num_threads=1,
    filter_vorticity=False,
    poisson_solver_type="greens_function_convolution",
    time=0.0,
    **kwargs,
):
    """``flow_simulators_3d.py:8-65``."""
    if flow_type == "navier_stokes":
        with_forcing = False
    elif flow_type == "navier_stokes_with_forcing":
        with_forcing = True
    else:
        msg = "Invalid flow type given"
        raise ValueError(msg)
    return UnboundedNavierStokesFlowSimulator3D(
        grid_size=grid_size,
        x_range=x_range,
        kinematic_viscosity=kinematic_viscosity,
        cfl=cfl,
        real_t=real_t,
        num_threads=num_threads,
        time=time,
        with_forcing=with_forcing,
        filter_vorticity=filter_vorticity,
        poisson_solver_type=poisson_solver_type,
        **kwargs,
    )
