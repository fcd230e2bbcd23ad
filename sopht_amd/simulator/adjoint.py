"""Differentiable coupled rollout: body forces of an immersed body as a function of its
kinematics, with gradients from hand-written adjoint kernels behind ``torch.autograd.Function``.

The reference has no differentiation. This module differentiates its driver loop
(``examples/3d_examples/FlowPastSphereCase/flow_past_sphere_case.py:191-197``: CFL ``dt``,
``interactor.time_step(dt)``, ``interactor()``, ``flow_sim.time_step(dt, U_inf)``) with
respect to the body's node positions and velocities, in reverse mode:

* the forward pass is that loop, run with the drop-in simulator and interactor; it keeps a
  checkpoint of the flow and virtual-boundary state every ``checkpoint_every`` steps;
* the backward pass recomputes one segment at a time from its checkpoint (out-of-place
  steps, ``sopht_ns3d_trajectory_step``), then walks it backwards with
  ``sopht_ns3d_adjoint_step`` (transposed stencils, the self-adjoint Green's convolution)
  and ``sopht_vbf_adjoint_3d`` (spreading and interpolation as each other's transposes,
  plus the weight gradient dw/dX of the cosine kernel);
* ``dt`` is frozen: the CFL maximum is not differentiated. A run may also be given a fixed
  ``dt_schedule`` (the finite-difference check uses the base run's).

The kinematics map parameters ``theta`` to node positions and velocities with torch
operations, so any torch-differentiable body motion works; :class:`HeaveKinematics` is the
rigid heave of a Fourier series.
"""

from __future__ import annotations

import ctypes as C
import math

import numpy as np

from .. import _lib as L


class HeaveKinematics:
    """Rigid heave in y: ``v(t) = sum_{m=1}^{M} a_m cos(2 pi m t / T) + b_m sin(2 pi m t / T)``,
    ``Y(t) = int_0^t v``, ``theta = (a_1..a_M, b_1..b_M)``; nodes ``X0 + (0, Y(t), 0)`` with
    velocity ``(0, v(t), 0)``."""

    def __init__(self, X0, period, n_modes=8, device=None):
        import torch

        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.X0 = torch.as_tensor(np.asarray(X0, dtype=np.float64)).to(self.device)
        self.period = float(period)
        self.n_modes = int(n_modes)
        self.num_params = 2 * self.n_modes
        self._host = None
        self._X = torch.empty_like(self.X0)
        self._V = torch.empty_like(self.X0)

    def prepare(self, theta):
        """Keep a host copy of ``theta`` for :meth:`state` (one synchronisation per run)."""
        self._host = (theta, theta.detach().to(dtype=__import__("torch").float64).cpu().numpy())

    def _basis_host(self, t):
        m = np.arange(1, self.n_modes + 1, dtype=np.float64)
        om = 2 * math.pi * m / self.period
        ph = om * float(t)
        return (np.concatenate([np.cos(ph), np.sin(ph)]),
                np.concatenate([np.sin(ph) / om, (1 - np.cos(ph)) / om]))

    def _basis(self, times):
        import torch

        t = torch.as_tensor(np.asarray(times, dtype=np.float64)).to(self.device).reshape(-1, 1)
        m = torch.arange(1, self.n_modes + 1, dtype=torch.float64, device=self.device)
        om = 2 * math.pi * m / self.period
        ph = om * t
        cv = torch.cat([torch.cos(ph), torch.sin(ph)], dim=1)
        cy = torch.cat([torch.sin(ph) / om, (1 - torch.cos(ph)) / om], dim=1)
        return cv, cy

    def trajectory(self, theta, times):
        """Positions and velocities ``(len(times), 3, n)`` float64 at the given times."""
        import torch

        cv, cy = self._basis(times)
        th = theta.to(torch.float64)
        Y = cy @ th
        v = cv @ th
        k = len(times)
        X = self.X0.unsqueeze(0).expand(k, -1, -1).clone()
        X[:, 1, :] = X[:, 1, :] + Y.reshape(k, 1)
        V = torch.zeros_like(X)
        V[:, 1, :] = v.reshape(k, 1)
        return X, V

    def state(self, theta, t):
        """Positions and velocities ``(3, n)`` float64 at time t. After :meth:`prepare` with the
        same ``theta`` the two scalars are formed on the host and the device work is four
        small kernels (no synchronisation); the returned tensors are reused by the next call."""
        if self._host is not None and self._host[0] is theta:
            cv, cy = self._basis_host(t)
            th = self._host[1]
            self._X.copy_(self.X0)
            self._X[1].add_(float(cy @ th))
            self._V.zero_()
            self._V[1].fill_(float(cv @ th))
            return self._X, self._V
        X, V = self.trajectory(theta, [t])
        return X[0], V[0]


class CoupledFlowRollout:
    """``num_steps`` steps of the coupled flow from the current state of ``flow`` and
    ``interactor`` (the state at construction is the start of every run)."""

    def __init__(self, flow, interactor, kinematics, free_stream_velocity, num_steps,
                 checkpoint_every=20):
        import torch

        if not flow.with_forcing:
            raise ValueError("the rollout needs a simulator built with forcing")
        if getattr(flow, "filter_vorticity", False):
            raise NotImplementedError("the adjoint does not cover the vorticity filter")
        if getattr(flow, "poisson_solver_type", "greens_function_convolution") != "greens_function_convolution":
            raise NotImplementedError("the adjoint covers the greens-function Poisson solver only")
        if interactor._ktype != 0:
            raise NotImplementedError("the adjoint supports the cosine interpolation kernel")
        self.flow = flow
        self.interactor = interactor
        self.kin = kinematics
        self.U = np.asarray(free_stream_velocity, dtype=np.float64).reshape(3)
        self.num_steps = int(num_steps)
        self.k = max(1, int(checkpoint_every))
        self.real_t = flow.real_t
        self._tag = L.dtype_tag(self.real_t)
        self.n = interactor.num_lag_nodes
        self.time0 = float(flow.time)
        self.initial = (
            flow.vorticity_field.clone(),
            flow.velocity_field.clone(),
            interactor.lag_grid_position_mismatch_field.clone(),
            interactor.lag_grid_velocity_mismatch_field.clone(),
        )
        self._dev = flow.vorticity_field.device
        self._tdt = flow.vorticity_field.dtype
        self._lib = L.lib()
        self.dts = None
        self.times = None
        self._shadow = None
        self._seg_ws = []  # one spreading workspace per step of a segment (see gradient)
        # One spreading workspace per step of the rollout when they fit the budget: the
        # recorded forward spreads step s through workspace s, and the backward's recompute of
        # step s (same node positions) then replays that plan instead of rebuilding it.
        self._step_ws = []
        self._recorded_with_step_ws = False
        if self.num_steps * self._ws_bytes() <= self.STEP_WS_BUDGET:  # allocated once, here
            self._workspace(self._step_ws, self.num_steps - 1)
        del torch

    # device bytes of one spreading workspace for n nodes, and the budget of the per-step set
    STEP_WS_BUDGET = 8 << 30

    def _ws_bytes(self):
        n = max(self.n, 1)
        padded = 1 << (n - 1).bit_length()
        return 4 * 64 * n * 4 + 2 * 8 * padded + 4 * (32768 + 1) + 64

    def _workspace(self, store, i):
        while len(store) <= i:
            w = C.c_void_p()
            L.call("sopht_spread3d_create", int(max(self.n, 1)), C.byref(w))
            store.append(w)
        return store[i]

    def __del__(self):
        for store in (getattr(self, "_seg_ws", []), getattr(self, "_step_ws", [])):
            for w in store:
                if w.value:
                    self._lib.sopht_spread3d_destroy(w)
        self._seg_ws = []
        self._step_ws = []

    # ---------------------------------------------------------------- forward
    def _restore(self):
        w0, u0, dX0, dV0 = self.initial
        self.flow.vorticity_field.copy_(w0)
        self.flow.velocity_field.copy_(u0)
        self.interactor.lag_grid_position_mismatch_field.copy_(dX0)
        self.interactor.lag_grid_velocity_mismatch_field.copy_(dV0)
        self.flow.eul_grid_forcing_field.zero_()
        self.flow.time = self.time0
        self.interactor.time = self.time0

    def run(self, theta, dt_schedule=None, record=True):
        """The driver loop; returns the body force of every step ``(num_steps, 3)`` (real_t,
        device). With ``record`` it keeps what the backward pass needs."""
        import torch

        flow, inter = self.flow, self.interactor
        self._restore()
        ns, k = self.num_steps, self.k
        forces = torch.empty((ns, 3), dtype=self._tdt, device=self._dev)
        dts = np.empty(ns)
        times = np.empty(ns)
        ck = []
        Xs = torch.empty((ns, 3, self.n), dtype=torch.float64, device=self._dev) if record else None
        Vs = torch.empty_like(Xs) if record else None
        theta = theta.detach()
        prep = getattr(self.kin, "prepare", None)
        if prep is not None:
            prep(theta)
        step_ws = record and len(self._step_ws) >= ns
        comm = inter.eul_lag_grid_communicator
        own_ws = comm._spread_ws
        t = self.time0
        for s in range(ns):
            if record and s % k == 0:
                ck.append(tuple(x.clone() for x in (
                    flow.vorticity_field, flow.velocity_field,
                    inter.lag_grid_position_mismatch_field, inter.lag_grid_velocity_mismatch_field)))
            # the body's state depends on t only: queued before the dt read-back waits
            X, V = self.kin.state(theta, t)
            if record:
                Xs[s].copy_(X)
                Vs[s].copy_(V)
            dt = float(dt_schedule[s]) if dt_schedule is not None else float(flow.compute_stable_timestep())
            dts[s] = dt
            times[s] = t
            inter.time_step(dt=dt)
            if step_ws:
                comm._spread_ws = self._workspace(self._step_ws, s)
            try:
                inter.compute_interaction_force_on_eul_and_lag_grid(
                    flow.eul_grid_forcing_field, flow.velocity_field, X, V)
            finally:
                comm._spread_ws = own_ws
            torch.sum(inter.lag_grid_forcing_field, dim=1, out=forces[s])
            forces[s].neg_()
            flow.time_step(dt=dt, free_stream_velocity=self.U)
            t += dt
        self.dts = dts
        self.times = times
        if record:
            self._ck = ck
            self._X = Xs
            self._V = Vs
            self._theta = theta.clone()
            self._recorded_with_step_ws = step_ws
        return forces

    # ---------------------------------------------------------------- backward
    def _shadow_interactor(self):
        if self._shadow is None:
            from ..numeric.immersed_boundary_ops import VirtualBoundaryForcing

            i = self.interactor
            self._shadow = VirtualBoundaryForcing(
                i.virtual_boundary_stiffness_coeff, i.virtual_boundary_damping_coeff, 3, i.dx,
                i.num_lag_nodes, i.real_t, eul_grid_coord_shift=i.eul_grid_coord_shift,
                enable_eul_grid_forcing_reset=False)
        return self._shadow

    def gradient(self, fbar):
        """``d(sum_s fbar_s . f_s)/d theta`` (float64) for the last recorded run."""
        import torch

        flow = self.flow
        ns, k, n = self.num_steps, self.k, self.n
        nz, ny, nx = flow.grid_size
        shape = (3, nz, ny, nx)
        dev, tdt = self._dev, self._tdt
        fbar = fbar.detach().to(device=dev, dtype=tdt).contiguous()
        sh = self._shadow_interactor()
        seg = min(k, ns)
        # per step of a segment: omega1 (the adjoint step's input) and u; omega ping-pongs and
        # F is scratch (the adjoint needs only omega1 and u)
        W = torch.empty((2, *shape), dtype=tdt, device=dev)
        W1 = torch.empty((seg, *shape), dtype=tdt, device=dev)
        U = torch.empty((seg + 1, *shape), dtype=tdt, device=dev)
        F = torch.empty(shape, dtype=tdt, device=dev)
        Fl = torch.empty((seg, 3, n), dtype=tdt, device=dev)
        wbar = torch.zeros(shape, dtype=tdt, device=dev)
        ubar = torch.zeros_like(wbar)
        work = torch.empty(shape, dtype=tdt, device=dev)
        dXb = torch.zeros((3, n), dtype=tdt, device=dev)
        dVb = torch.zeros_like(dXb)
        Xb = torch.empty((ns, 3, n), dtype=tdt, device=dev)
        Vb = torch.empty_like(Xb)
        wts = torch.empty((64, n), dtype=tdt, device=dev)
        idx = torch.empty((3, n), dtype=torch.int64, device=dev)
        lag = torch.empty((3, n), dtype=tdt, device=dev)
        dx = float(flow.dx)
        shift = float(sh.eul_grid_coord_shift)
        kk = float(sh.virtual_boundary_stiffness_coeff)
        cc = float(sh.virtual_boundary_damping_coeff)
        st = L.stream_ptr()
        lib = self._lib
        # step s spreads with the workspace the recorded forward used for it (its plan is then
        # replayed by the recompute and by the reverse sweep), or else with workspace j of the
        # segment, whose plan the recompute builds and the reverse sweep replays
        if self._recorded_with_step_ws:
            ws_of = lambda s, j: self._step_ws[s]  # noqa: E731
        else:
            for j in range(seg):
                self._workspace(self._seg_ws, j)
            ws_of = lambda s, j: self._seg_ws[j]  # noqa: E731
        Fst = L.strides3(F[0])
        Fp = L.comps(F)
        for g in reversed(range(len(self._ck))):
            s0 = g * k
            s1 = min(ns, s0 + k)
            w0, u0, dX0, dV0 = self._ck[g]
            W[0].copy_(w0)
            U[0].copy_(u0)
            sh.lag_grid_position_mismatch_field.copy_(dX0)
            sh.lag_grid_velocity_mismatch_field.copy_(dV0)
            for j, s in enumerate(range(s0, s1)):
                dt = float(self.dts[s])
                sh.time_step(dt=dt)
                F.zero_()
                sh.compute_interaction_force_on_lag_grid(U[j], self._X[s], self._V[s])
                Fl[j].copy_(sh.lag_grid_forcing_field)
                L.check(lib.sopht_lag_to_eul_spread_3d(
                    self._tag, 3, ws_of(s, j), Fp, Fst, nz, ny, nx, L.ptr(Fl[j]),
                    L.ptr(sh.interp_weights), L.ptr(sh.nearest_eul_grid_index_to_lag_grid), n, st))
                flow.trajectory_step(W[j % 2], U[j], F, W[(j + 1) % 2], U[j + 1], dt, self.U,
                                     vorticity1_out=W1[j])
            for j in reversed(range(s1 - s0)):
                s = s0 + j
                dt = float(self.dts[s])
                L.check(lib.sopht_ns3d_adjoint_step(
                    flow._plan, L.ptr(W1[j]), L.ptr(U[j]), dt, L.ptr(wbar), L.ptr(ubar),
                    L.ptr(work), st))
                pF = dt / (2.0 * dx * float(flow.flow_density))
                L.check(lib.sopht_vbf_adjoint_3d(
                    self._tag, L.ptr(wbar), pF, L.ptr(U[j]), nz, ny, nx, L.ptr(self._X[s]),
                    L.ptr(Fl[j]), n, dx, shift, kk, cc, dt, L.ptr(fbar[s]), L.ptr(dXb), L.ptr(dVb),
                    L.ptr(Xb[s]), L.ptr(Vb[s]), L.ptr(wts), L.ptr(idx), L.ptr(lag), ws_of(s, j),
                    L.ptr(ubar), st))
        with torch.enable_grad():
            th = self._theta.detach().clone().requires_grad_(True)
            Xt, Vt = self.kin.trajectory(th, self.times)
            torch.autograd.backward([Xt, Vt], [Xb.to(torch.float64), Vb.to(torch.float64)])
        self.initial_vorticity_adjoint = wbar
        self.initial_velocity_adjoint = ubar
        return th.grad.detach()


class IBFlowRolloutFunction:
    """``torch.autograd.Function``: ``theta -> body forces (num_steps, 3)`` of a
    :class:`CoupledFlowRollout`; backward runs the checkpointed adjoint."""

    _fn = None

    @classmethod
    def get(cls):
        if cls._fn is None:
            import torch

            class _Fn(torch.autograd.Function):
                @staticmethod
                def forward(ctx, theta, rollout, dt_schedule):
                    forces = rollout.run(theta, dt_schedule=dt_schedule, record=True)
                    ctx.rollout = rollout
                    return forces

                @staticmethod
                def backward(ctx, fbar):
                    th = ctx.rollout.gradient(fbar)
                    return th, None, None

            cls._fn = _Fn
        return cls._fn

    @classmethod
    def apply(cls, theta, rollout, dt_schedule=None):
        return cls.get().apply(theta, rollout, dt_schedule)


def body_forces(theta, rollout, dt_schedule=None):
    """Differentiable body forces ``(num_steps, 3)`` of ``rollout`` for parameters ``theta``."""
    return IBFlowRolloutFunction.apply(theta, rollout, dt_schedule)


def time_averaged_force(forces, dts, axis=0):
    """``sum_s f_s[axis] dt_s / sum_s dt_s`` (dt frozen)."""
    import torch

    w = torch.as_tensor(np.asarray(dts, dtype=np.float64) / float(np.sum(dts))).to(forces.device)
    return (forces[:, axis].to(torch.float64) * w).sum()
