"""Drop-in ``EulerianLagrangianGridCommunicator3D`` on the gfx950 HIP kernels.

Mirrors ``sopht/numeric/immersed_boundary_ops/EulerianLagrangianGridCommunicator3D.py:7-65``:
the same four kernel attributes with the same positional arguments, operating on device
tensors:

* positions ``(3, n)`` float64, nearest index ``(3, n)`` int64 (x, y, z rows);
* support ``(3, 4, 4, 4, n)`` and weights ``(4, 4, 4, n)`` in real_t, block indexed
  ``[kz, ky, kx]``.

Spreading is deterministic (sort by support block, per-cell ordered gather) and adds in
the reference's node order; nodes whose support leaves the grid raise ``IndexError`` on the
next call that checks (the reference reads out of bounds there).
"""

from __future__ import annotations

import ctypes as C

from ... import _lib as L

_KERNEL_TYPES = {"cosine": 0, "peskin": 1}


class EulerianLagrangianGridCommunicator3D:
    """Eulerian <-> Lagrangian transfer with the 4^3 cosine or Peskin delta kernel."""

    def __init__(
        self,
        dx,
        eul_grid_coord_shift,
        num_lag_nodes,
        interp_kernel_width,
        real_t,
        n_components=1,
        interp_kernel_type="cosine",
    ):
        if n_components not in (1, 3):
            msg = "Invalid number of components for interpolation, must be either 1 or 3"
            raise ValueError(msg)
        if interp_kernel_type not in _KERNEL_TYPES:
            msg = (
                "Invalid interpolation kernel type. Current supported types are"
                "'cosine' and 'peskin'."
            )
            raise ValueError(msg)
        if interp_kernel_width != 2:
            msg = "Interpolation kernel inconsistent with interpolation kernel width!"
            raise ValueError(msg)
        self.dx = dx
        self.eul_grid_coord_shift = eul_grid_coord_shift
        self.num_lag_nodes = num_lag_nodes
        self.interp_kernel_width = interp_kernel_width
        self.real_t = real_t
        self.n_components = n_components
        self.interp_kernel_type = interp_kernel_type
        self._tag = L.dtype_tag(real_t)
        self._ktype = _KERNEL_TYPES[interp_kernel_type]
        ws = C.c_void_p()
        L.call("sopht_spread3d_create", int(num_lag_nodes), C.byref(ws))
        self._spread_ws = ws
        self._lib = L.lib()

        self.local_eulerian_grid_support_of_lagrangian_grid_kernel = self._support_kernel
        self.eulerian_to_lagrangian_grid_interpolation_kernel = self._interp_kernel
        self.lagrangian_to_eulerian_grid_interpolation_kernel = self._spread_kernel
        self.interpolation_weights_kernel = self._weights_kernel

    def __del__(self):
        ws = getattr(self, "_spread_ws", None)
        if ws is not None and ws.value:
            self._lib.sopht_spread3d_destroy(ws)
            self._spread_ws = None

    # :102-135
    def _support_kernel(
        self, local_eul_grid_support_of_lag_grid, nearest_eul_grid_index_to_lag_grid, lag_positions
    ):
        n = int(lag_positions.shape[-1])
        lag_positions = lag_positions.contiguous()
        L.call(
            "sopht_lag_support_3d", self._tag, L.ptr(lag_positions),
            L.ptr(nearest_eul_grid_index_to_lag_grid), L.ptr(local_eul_grid_support_of_lag_grid),
            n, float(self.dx), float(self.eul_grid_coord_shift), L.stream_ptr(),
        )
        L.written(nearest_eul_grid_index_to_lag_grid, local_eul_grid_support_of_lag_grid)

    # :401-410 (cosine) / :429-516 (peskin)
    def _weights_kernel(self, interp_weights, local_eul_grid_support_of_lag_grid):
        n = int(interp_weights.shape[-1])
        L.call(
            "sopht_lag_interp_weights_3d", self._tag, self._ktype, L.ptr(interp_weights),
            L.ptr(local_eul_grid_support_of_lag_grid), n, float(self.dx), L.stream_ptr(),
        )
        L.written(interp_weights, local_eul_grid_support_of_lag_grid)

    def _eul_ptrs(self, eul_grid_field):
        if eul_grid_field.dim() == 4:
            return L.comps(eul_grid_field), L.strides3(eul_grid_field), eul_grid_field.shape[1:], 3
        p = (C.c_void_p * 3)(eul_grid_field.data_ptr(), 0, 0)
        return p, L.strides3(eul_grid_field), eul_grid_field.shape, 1

    # :198-295
    def _interp_kernel(self, lag_grid_field, eul_grid_field, interp_weights, nearest_eul_grid_index_to_lag_grid):
        p, st, shp, nc = self._eul_ptrs(eul_grid_field)
        n = int(interp_weights.shape[-1])
        L.call(
            "sopht_eul_to_lag_interp_3d", self._tag, nc, L.ptr(lag_grid_field), p, st,
            int(shp[0]), int(shp[1]), int(shp[2]), L.ptr(interp_weights),
            L.ptr(nearest_eul_grid_index_to_lag_grid), n, float(self.dx), L.stream_ptr(),
        )
        L.written(lag_grid_field)

    # :318-380
    def _spread_kernel(self, eul_grid_field, lag_grid_field, interp_weights, nearest_eul_grid_index_to_lag_grid):
        p, st, shp, nc = self._eul_ptrs(eul_grid_field)
        n = int(interp_weights.shape[-1])
        L.call(
            "sopht_lag_to_eul_spread_3d", self._tag, nc, self._spread_ws, p, st, int(shp[0]),
            int(shp[1]), int(shp[2]), L.ptr(lag_grid_field.contiguous()), L.ptr(interp_weights),
            L.ptr(nearest_eul_grid_index_to_lag_grid), n, L.stream_ptr(),
        )
        L.written(eul_grid_field)

    def declare_support_unchanged(self):
        """No counterpart in the reference: the next spreading call has the node count and
        support indices of the previous one (a body at rest), so it replays the spreading plan
        without the sort and plan kernels. A support that changed anyway is caught on the
        device and reported by the next ``check_support_in_domain`` (``ValueError``)."""
        L.call("sopht_spread3d_reuse_plan", self._spread_ws)

    def check_support_in_domain(self):
        """Synchronise and raise ``IndexError`` if a spread support block left the grid."""
        L.call("sopht_spread3d_check", self._spread_ws, L.stream_ptr())
