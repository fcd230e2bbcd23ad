"""Drop-in ``VirtualBoundaryForcing`` (3D) on the gfx950 HIP kernels.

Mirrors ``sopht/numeric/immersed_boundary_ops/VirtualBoundaryForcing.py:20-283``: same
constructor, public buffers (device tensors of the reference shapes) and methods. The
interaction force on the Lagrangian grid (support, weights, interpolation, mismatch and
penalty force, ``:187-230``) runs as one fused kernel; spreading uses the communicator's
deterministic ordered gather. Lagrangian positions and body velocities may be given as
host numpy arrays (they are uploaded, float64) or as float64 device tensors.
"""

from __future__ import annotations

import numpy as np

from ... import _lib as L
from .eulerian_lagrangian_grid_communicator_3d import EulerianLagrangianGridCommunicator3D


def _device():
    import torch

    return torch.device("cuda", torch.cuda.current_device())


def as_device_f64(a, device=None):
    """float64 contiguous device tensor view/copy of a host or device ``(3, n)`` array."""
    import torch

    device = device or _device()
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=torch.float64).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).to(device, non_blocking=False)


class _LagrangianUpload:
    """Device copy of a host ``(3, n)`` float64 array, refreshed only when its values change.

    A rigid body at rest presents the same positions every step; re-sending them would
    cost a synchronous host-to-device copy per call. New values go through a pinned
    staging buffer with an asynchronous copy on the current stream; the staging buffer is
    rewritten only after the previous copy from it has completed."""

    def __init__(self):
        self.host = None
        self.dev = None
        self.pinned = None
        self.event = None
        self.generation = 0  # bumped whenever the values may have changed

    def get(self, a):
        import torch

        if isinstance(a, torch.Tensor):
            self.generation += 1  # a device tensor: its values are not tracked
            return as_device_f64(a)
        a = np.asarray(a, dtype=np.float64)
        if self.host is not None and self.host.shape == a.shape and np.array_equal(self.host, a):
            return self.dev
        self.generation += 1
        if self.dev is None or tuple(self.dev.shape) != a.shape:
            self.dev = torch.empty(a.shape, dtype=torch.float64, device=_device())
            self.pinned = torch.empty(a.shape, dtype=torch.float64, pin_memory=True)
            self.event = None
        if self.event is not None:
            self.event.synchronize()
        self.pinned.numpy()[...] = a
        self.dev.copy_(self.pinned, non_blocking=True)
        self.event = torch.cuda.Event()
        self.event.record()
        self.host = a.copy()
        return self.dev


class VirtualBoundaryForcing:
    """Goldstein (1993) virtual boundary feedback between body and flow."""

    def __init__(
        self,
        virtual_boundary_stiffness_coeff,
        virtual_boundary_damping_coeff,
        grid_dim,
        dx,
        num_lag_nodes,
        real_t,
        eul_grid_coord_shift=None,
        interp_kernel_width=None,
        enable_eul_grid_forcing_reset=True,
        num_threads=False,
        start_time=0.0,
        interp_kernel_type="cosine",
    ):
        import torch

        if grid_dim not in (2, 3):
            msg = "Invalid grid dimensions, must be either 2 or 3"
            raise ValueError(msg)
        if grid_dim == 2:
            msg = "2D virtual boundary forcing is outside this backend (3D flow step only)"
            raise NotImplementedError(msg)
        self.grid_dim = grid_dim
        self.virtual_boundary_stiffness_coeff = virtual_boundary_stiffness_coeff
        self.virtual_boundary_damping_coeff = virtual_boundary_damping_coeff
        self.time = start_time
        if eul_grid_coord_shift is None:
            eul_grid_coord_shift = real_t(dx / 2)
        if interp_kernel_width is None:
            interp_kernel_width = 2
        self.dx = dx
        self.real_t = real_t
        self.eul_grid_coord_shift = eul_grid_coord_shift
        self.num_lag_nodes = num_lag_nodes
        dev = _device()
        tdt = L.torch_dtype(real_t)
        n = num_lag_nodes
        k = 2 * interp_kernel_width
        self.nearest_eul_grid_index_to_lag_grid = torch.empty((3, n), dtype=torch.int64, device=dev)
        self.local_eul_grid_support_of_lag_grid = torch.empty((3, k, k, k, n), dtype=tdt, device=dev)
        self.interp_weights = torch.empty((k, k, k, n), dtype=tdt, device=dev)
        self.lag_grid_flow_velocity_field = torch.zeros((3, n), dtype=tdt, device=dev)
        self.lag_grid_position_mismatch_field = torch.zeros_like(self.lag_grid_flow_velocity_field)
        self.lag_grid_velocity_mismatch_field = torch.zeros_like(self.lag_grid_flow_velocity_field)
        self.lag_grid_forcing_field = torch.zeros_like(self.lag_grid_flow_velocity_field)
        self.eul_lag_grid_communicator = EulerianLagrangianGridCommunicator3D(
            dx=dx,
            eul_grid_coord_shift=eul_grid_coord_shift,
            num_lag_nodes=n,
            interp_kernel_width=interp_kernel_width,
            real_t=real_t,
            n_components=3,
            interp_kernel_type=interp_kernel_type,
        )
        self._ktype = self.eul_lag_grid_communicator._ktype
        self._tag = L.dtype_tag(real_t)
        self._pos_upload = _LagrangianUpload()
        self._vel_upload = _LagrangianUpload()
        # (position generation, workspace) of the last spreading: the same positions through
        # the same workspace have the same support indices, and its plan is replayed without
        # the sort and plan kernels
        self._spread_key = None
        self.enable_eul_grid_forcing_reset = enable_eul_grid_forcing_reset
        if enable_eul_grid_forcing_reset:
            self.compute_interaction_forcing = (
                self.compute_interaction_force_on_eul_and_lag_grid_with_eul_grid_forcing_reset
            )
        else:
            self.compute_interaction_forcing = self.compute_interaction_force_on_eul_and_lag_grid

    # :132-147
    @staticmethod
    def compute_lag_grid_velocity_mismatch_field(
        lag_grid_velocity_mismatch_field, lag_grid_flow_velocity_field, lag_grid_body_velocity_field
    ):
        lag_grid_velocity_mismatch_field.copy_(
            (lag_grid_flow_velocity_field.double() - as_device_f64(lag_grid_body_velocity_field))
        )

    # :149-164
    @staticmethod
    def update_lag_grid_position_mismatch_field_via_euler_forward(
        lag_grid_position_mismatch_field, lag_grid_velocity_mismatch_field, dt
    ):
        t = lag_grid_position_mismatch_field
        L.call(
            "sopht_vbf_time_step_3d", L.dtype_tag(np.float32 if t.dtype.itemsize == 4 else np.float64),
            L.ptr(t), L.ptr(lag_grid_velocity_mismatch_field), int(t.shape[-1]), float(dt),
            L.stream_ptr(),
        )
        L.written(t)

    # :166-185
    @staticmethod
    def compute_lag_grid_forcing_field(
        lag_grid_forcing_field,
        lag_grid_position_mismatch_field,
        lag_grid_velocity_mismatch_field,
        virtual_boundary_stiffness_coeff,
        virtual_boundary_damping_coeff,
    ):
        lag_grid_forcing_field.copy_(
            virtual_boundary_stiffness_coeff * lag_grid_position_mismatch_field.double()
            + virtual_boundary_damping_coeff * lag_grid_velocity_mismatch_field.double()
        )

    # :187-230, fused
    def compute_interaction_force_on_lag_grid(
        self, eul_grid_velocity_field, lag_grid_position_field, lag_grid_velocity_field
    ):
        L.require_device_tensor(eul_grid_velocity_field, self.real_t, "eul_grid_velocity_field", 4)
        X = self._pos_upload.get(lag_grid_position_field)
        V = self._vel_upload.get(lag_grid_velocity_field)
        u = eul_grid_velocity_field
        nz, ny, nx = (int(s) for s in u.shape[1:])
        L.call(
            "sopht_vbf_lag_force_3d", self._tag, self._ktype, L.comps(u), L.strides3(u), nz, ny, nx,
            L.ptr(X), L.ptr(V), self.num_lag_nodes, float(self.dx), float(self.eul_grid_coord_shift),
            float(self.virtual_boundary_stiffness_coeff), float(self.virtual_boundary_damping_coeff),
            L.ptr(self.nearest_eul_grid_index_to_lag_grid), L.ptr(self.local_eul_grid_support_of_lag_grid),
            L.ptr(self.interp_weights), L.ptr(self.lag_grid_flow_velocity_field),
            L.ptr(self.lag_grid_velocity_mismatch_field), L.ptr(self.lag_grid_position_mismatch_field),
            L.ptr(self.lag_grid_forcing_field), L.stream_ptr(),
        )
        L.written(
            self.nearest_eul_grid_index_to_lag_grid, self.local_eul_grid_support_of_lag_grid,
            self.interp_weights, self.lag_grid_flow_velocity_field,
            self.lag_grid_velocity_mismatch_field, self.lag_grid_position_mismatch_field,
            self.lag_grid_forcing_field,
        )

    # :232-253
    def compute_interaction_force_on_eul_and_lag_grid(
        self, eul_grid_forcing_field, eul_grid_velocity_field, lag_grid_position_field, lag_grid_velocity_field
    ):
        self.compute_interaction_force_on_lag_grid(
            eul_grid_velocity_field, lag_grid_position_field, lag_grid_velocity_field
        )
        comm = self.eul_lag_grid_communicator
        key = (self._pos_upload.generation, comm._spread_ws.value)
        if key == self._spread_key:
            comm.declare_support_unchanged()
        self._spread_key = key
        self.eul_lag_grid_communicator.lagrangian_to_eulerian_grid_interpolation_kernel(
            eul_grid_forcing_field,
            self.lag_grid_forcing_field,
            self.interp_weights,
            self.nearest_eul_grid_index_to_lag_grid,
        )

    # :255-274
    def compute_interaction_force_on_eul_and_lag_grid_with_eul_grid_forcing_reset(
        self, eul_grid_forcing_field, eul_grid_velocity_field, lag_grid_position_field, lag_grid_velocity_field
    ):
        eul_grid_forcing_field.zero_()
        self.compute_interaction_force_on_eul_and_lag_grid(
            eul_grid_forcing_field, eul_grid_velocity_field, lag_grid_position_field, lag_grid_velocity_field
        )

    # :276-283
    def time_step(self, dt):
        self.update_lag_grid_position_mismatch_field_via_euler_forward(
            self.lag_grid_position_mismatch_field, self.lag_grid_velocity_mismatch_field, dt
        )
        self.time += dt

    def check_support_in_domain(self):
        """Synchronise; ``IndexError`` if a node's support left the Eulerian grid."""
        self.eul_lag_grid_communicator.check_support_in_domain()
