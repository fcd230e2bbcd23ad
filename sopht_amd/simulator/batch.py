"""Many independent coupled flow simulations advanced together on one GPU.

Each member is an ``UnboundedNavierStokesFlowSimulator3D`` (with forcing) plus, optionally,
the ``ImmersedBodyFlowInteraction`` that drives its forcing field, and a free-stream
velocity. One batched step of a member is the reference driver's loop body
(``flow_past_sphere_case.py:191-197``)::

    dt = flow.compute_stable_timestep()
    interactor.time_step(dt)
    interactor()
    flow.time_step(dt, free_stream_velocity=U)

with dt computed on the device, so the whole step runs without a host round trip and is
captured once into a hipGraph and replayed. All members share one launch per stage (the
Poisson passes run once per distinct grid size). Member results are bit-identical to
stepping each member alone with that loop.

Limits of this backend: a member's body keeps the Lagrangian positions and velocities it
has when the batch is built (rigid bodies held in place or in uniform motion of their
forcing grid), at most 8192 Lagrangian nodes per member, and ``flow.time`` and the
interactor's own time are updated on :meth:`synchronize`.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib as L


class BatchedFlowSimulation:
    """Advance ``members`` = [(flow, interactor_or_None, free_stream_velocity), ...] together."""

    def __init__(self, members):
        import torch

        if not members:
            msg = "a batch needs at least one member"
            raise ValueError(msg)
        self._flows = [m[0] for m in members]
        self._inters = [m[1] for m in members]
        real_t = self._flows[0].real_t
        if any(f.real_t != real_t for f in self._flows):
            msg = "all members of a batch share one real_t"
            raise ValueError(msg)
        self.real_t = real_t
        self._keep = []  # device copies the batch reads (positions, velocities)
        arr = (L.BatchMember * len(members))()
        for i, (flow, inter, U) in enumerate(members):
            if not flow.with_forcing and inter is not None:
                msg = "a member with an interactor needs with_forcing=True"
                raise ValueError(msg)
            if getattr(flow, "filter_vorticity", False):
                msg = "the batched step does not cover the vorticity filter"
                raise NotImplementedError(msg)
            if getattr(flow, "poisson_solver_type", "greens_function_convolution") != "greens_function_convolution":
                msg = "the batched step covers the greens-function Poisson solver only"
                raise NotImplementedError(msg)
            q = arr[i]
            nz, ny, nx = (int(v) for v in flow.grid_size)
            q.nz, q.ny, q.nx = nz, ny, nx
            q.x_range = float(flow.x_range)
            q.nu = float(flow.kinematic_viscosity)
            q.rho = float(flow.flow_density)
            q.cfl = float(flow.cfl)
            q.width = int(flow.penalty_zone_width)
            q.with_forcing = int(bool(flow.with_forcing))
            q.with_free_stream = int(bool(flow.with_free_stream_flow))
            U = np.zeros(3) if U is None else np.asarray(U, dtype=np.float64).reshape(3)
            q.free_stream[:] = [float(v) for v in U]
            q.time = float(flow.time)
            q.vorticity = flow.vorticity_field.data_ptr()
            q.velocity = flow.velocity_field.data_ptr()
            q.psi = flow.stream_func_field.data_ptr()
            q.forcing = flow.eul_grid_forcing_field.data_ptr() if flow.with_forcing else None
            q.n_lag = 0
            if inter is not None:
                grid = inter.forcing_grid
                grid.compute_lag_grid_position_field()
                grid.compute_lag_grid_velocity_field()
                X = torch.as_tensor(np.ascontiguousarray(grid.position_field, dtype=np.float64)).cuda()
                V = torch.as_tensor(np.ascontiguousarray(grid.velocity_field, dtype=np.float64)).cuda()
                self._keep += [X, V]
                q.n_lag = int(inter.num_lag_nodes)
                q.kernel_type = int(inter._ktype)
                q.dx = float(inter.dx)
                q.shift = float(inter.eul_grid_coord_shift)
                q.k = float(inter.virtual_boundary_stiffness_coeff)
                q.c = float(inter.virtual_boundary_damping_coeff)
                q.X, q.V = X.data_ptr(), V.data_ptr()
                q.idx = inter.nearest_eul_grid_index_to_lag_grid.data_ptr()
                q.support = inter.local_eul_grid_support_of_lag_grid.data_ptr()
                q.weights = inter.interp_weights.data_ptr()
                q.lag_u = inter.lag_grid_flow_velocity_field.data_ptr()
                q.dV = inter.lag_grid_velocity_mismatch_field.data_ptr()
                q.dX = inter.lag_grid_position_mismatch_field.data_ptr()
                q.Fl = inter.lag_grid_forcing_field.data_ptr()
        self._members = arr
        handle = C.c_void_p()
        L.call("sopht_ns3d_batch_create", L.dtype_tag(real_t), len(members), arr, L.stream_ptr(), C.byref(handle))
        self._batch = handle
        self._time0 = [float(f.time) for f in self._flows]

    def step(self, steps: int = 1, use_graph: bool = True) -> None:
        """Advance every member by ``steps`` steps (asynchronous on the current stream)."""
        L.call("sopht_ns3d_batch_step", self._batch, int(steps), int(bool(use_graph)), L.stream_ptr())
        # the members' fields were written through raw pointers: their version counters
        # must move, or a member's cached velocity maxima would still look current
        for flow, inter in zip(self._flows, self._inters):
            L.written(flow.vorticity_field, flow.velocity_field, flow.stream_func_field,
                      flow.eul_grid_forcing_field if flow.with_forcing else None)
            if inter is not None:
                L.written(inter.lag_grid_flow_velocity_field, inter.lag_grid_velocity_mismatch_field,
                          inter.lag_grid_position_mismatch_field, inter.lag_grid_forcing_field)

    def poisson_schedules(self) -> list:
        """Per size class: its grid, component count and its Poisson plan's FFT schedule
        (passes, z-pass jobs per workgroup chosen when the plan was created)."""
        out = []
        lib = L.lib()
        n = lib.sopht_ns3d_batch_class_schedule(self._batch, -1, None, None, None, None)
        L.check(min(0, n))
        for i in range(n):
            shape = (C.c_int64 * 3)()
            nc, passes, jobs = C.c_int(0), C.c_int(0), C.c_int(0)
            L.check(min(0, lib.sopht_ns3d_batch_class_schedule(
                self._batch, i, shape, C.byref(nc), C.byref(passes), C.byref(jobs))))
            out.append({"grid": list(shape), "components": nc.value, "passes": passes.value,
                        "zpass_jobs": jobs.value})
        return out

    def synchronize(self):
        """Wait for the device; set each member's ``time`` (and its interactor's) and return
        the last dt of every member."""
        n = len(self._flows)
        dt = (C.c_double * n)()
        t = (C.c_double * n)()
        L.call("sopht_ns3d_batch_state", self._batch, dt, t, L.stream_ptr())
        for i, flow in enumerate(self._flows):
            flow.time = self.real_t(t[i])
            if self._inters[i] is not None:
                self._inters[i].time = flow.time
        return np.array([dt[i] for i in range(n)])

    def close(self) -> None:
        if getattr(self, "_batch", None):
            L.lib().sopht_ns3d_batch_destroy(self._batch)
            self._batch = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
