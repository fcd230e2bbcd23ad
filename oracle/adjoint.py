"""Oracle: discrete adjoint of the coupled IB flow rollout (test only).

TEST INFRASTRUCTURE ONLY (see the package header). The reference has no adjoint; this
module differentiates the reference's own discrete step (the driver loop of
``examples/3d_examples/FlowPastSphereCase/flow_past_sphere_case.py:191-197``) in reverse
mode, with the time step frozen (``dt`` comes from a non-differentiable maximum):

per step n (state omega, u, dX, dV_prev; dt_n; node positions X_n and velocities V_n)
  dX  <- dX + dt dV_prev                                  VirtualBoundaryForcing.py:149-164
  w    = W(X_n); U = dx^3 I(u; w); dV = U - V_n; Fl = k dX + c dV     :187-230
  f_n  = -sum_nodes Fl                                    rigid_body_forcing_grids.py:152-169
  F    = S(Fl; w)                                         EulerianLagrangianGridCommunicator3D.py:318-380
  w1   = omega + pF M C F                                 navier_stokes_flow_simulators.py:487-498
  w2   = w1 + pB M C (u x w1)                             :454-465
  w3   = w2 + pD M L w2                                   :466-471
  omega' = P w3                                           :474
  u'   = pC M C G omega' + U_inf                          :475-485

C is the unnormalised central curl, L the 7-point Laplacian, M the interior mask of the
radius-1 stencils (writes only on [1:-1]^3), P the clamp-and-sine penalisation, G the
free-space Green's convolution, I / S interpolation and spreading. Their transposes:
C^T = C (evaluated on every cell, neighbours outside the grid are zero), L^T = L in the
same sense, G^T = G (the doubled-domain kernel is even), S^T = I / dx^3, P^T a gather of
the border slabs onto the clamp planes. The weight derivative dw/dX of the cosine kernel
carries the body kinematics.

Pinning (tests/test_oracle_adjoint.py): each transpose satisfies <A x, y> = <x, A^T y>
against the forward operators of ``oracle.stencils`` / ``oracle.ib`` / ``oracle.poisson``
on random inputs; dw/dX matches central differences of the weights; the rollout gradient
matches central finite differences of the rollout (frozen dt) in FP64.
"""

from __future__ import annotations

import numpy as np

from . import flow as flow_mod
from . import ib, stencils

# ----------------------------------------------------------------- linear transposes


def _masked(a):
    """``a`` with its radius-1 border set to zero (the interior mask M)."""
    m = np.zeros_like(a)
    m[..., 1:-1, 1:-1, 1:-1] = a[..., 1:-1, 1:-1, 1:-1]
    return m


def _shift(a, axis, off):
    """``a`` shifted so that out[q] = a[q + off] along spatial ``axis`` (0=z, 1=y, 2=x),
    zero outside the grid."""
    out = np.zeros_like(a)
    n = a.shape[axis]
    src = [slice(None)] * 3
    dst = [slice(None)] * 3
    if off > 0:
        src[axis] = slice(off, n)
        dst[axis] = slice(0, n - off)
    else:
        src[axis] = slice(0, n + off)
        dst[axis] = slice(-off, n)
    out[tuple(dst)] = a[tuple(src)]
    return out


def curl_full(a, prefactor):
    """``prefactor * curl_h(a)`` on every cell with zero outside the grid (no mask)."""
    ax, ay, az = a[0], a[1], a[2]
    d = _shift
    cx = (d(az, 1, 1) - d(az, 1, -1)) - (d(ay, 0, 1) - d(ay, 0, -1))
    cy = (d(ax, 0, 1) - d(ax, 0, -1)) - (d(az, 2, 1) - d(az, 2, -1))
    cz = (d(ay, 2, 1) - d(ay, 2, -1)) - (d(ax, 1, 1) - d(ax, 1, -1))
    return prefactor * np.stack([cx, cy, cz])


def curl_transpose(g, prefactor):
    """Transpose of ``f -> prefactor * M curl_h(f)`` (interior-only curl, ``curl_3d.py:37-66``):
    ``prefactor * curl(M g)`` on every cell."""
    return curl_full(_masked(g), prefactor)


def laplacian_full(a):
    """7-point Laplacian on every cell with zero outside the grid."""
    out = -6.0 * a
    for axis in range(3):
        out = out + _shift(a, axis, 1) + _shift(a, axis, -1)
    return out


def diffusion_transpose(g, pD):
    """Transpose of ``w -> w + pD M L w`` per component (``diffusion_flux_3d.py:36-44``,
    ``diffusion_timestep_3d.py:44``): ``g + pD L(M g)``."""
    return g + pD * np.stack([laplacian_full(c) for c in _masked(g)])


def penalise_axis_matrix(n, width, dx, grid_1d):
    """Matrix P_a (n x n) of one axis of the penalisation: ``(P_a f)[p] = s(p) f[clamp(p)]``
    with ``clamp`` to [w-1, n-w] and ``s`` the sine factor of ``penalise_field_boundary_3d.py``
    (1 off the border slabs)."""
    P = np.zeros((n, n))
    if width == 0:
        return np.eye(n)
    w = width
    sp = (np.pi / 2) / (w * dx)
    for p in range(n):
        c = min(max(p, w - 1), n - w)
        s = 1.0
        if p < w:
            s = np.sin(sp * (grid_1d[p] - grid_1d[0]))
        elif p >= n - w:
            s = np.sin(sp * (grid_1d[-1] - grid_1d[p]))
        P[p, c] = s
    return P


def penalise_matrices(flow):
    """(P_z, P_y, P_x) of an oracle simulator."""
    pf = flow.position_field
    gx, gy, gz = pf[0][0, 0, :], pf[1][0, :, 0], pf[2][:, 0, 0]
    w = flow.penalty_zone_width
    return (
        penalise_axis_matrix(len(gz), w, flow.dx, gz),
        penalise_axis_matrix(len(gy), w, flow.dx, gy),
        penalise_axis_matrix(len(gx), w, flow.dx, gx),
    )


def penalise_apply(field, mats, transpose=False):
    Pz, Py, Px = (m.T if transpose else m for m in mats)
    return np.einsum("ai,bj,ck,...ijk->...abc", Pz, Py, Px, field, optimize=True)


def cross(a, b):
    """``a x b`` per cell with the reference's component formulas (``elementwise_ops_3d.py:427-447``)."""
    return np.stack(
        [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]
    )


# ----------------------------------------------------------------- weights and their gradient


def cosine_weight_gradient(idx, lag_positions, dx, shift):
    """``dw/dX_a`` of the cosine kernel (``EulerianLagrangianGridCommunicator3D.py:387-412``):
    ``w = (0.25/dx)^3 prod_b phi(s_b)``, ``phi(s) = 1 + cos(pi s / 2)``, ``s_b = d_b/dx``,
    ``d_b = (idx_b + off_b) dx + shift - X_b``, so
    ``dw/dX_a = (0.25/dx)^3 (pi/(2 dx)) sin(pi s_a/2) prod_{b != a} phi(s_b)``.
    Returns (3, 4, 4, 4, n) indexed [a, kz, ky, kx, node]."""
    n = lag_positions.shape[1]
    off = ib.support_offsets()
    oz, oy, ox = np.meshgrid(off, off, off, indexing="ij")
    offs = np.stack((ox, oy, oz)).reshape(3, 4, 4, 4, 1)
    d = (idx.reshape(3, 1, 1, 1, n) + offs) * dx + shift - lag_positions.reshape(3, 1, 1, 1, n)
    s = d / dx
    phi = 1.0 + np.cos(0.5 * np.pi * s)
    dphi = (0.5 * np.pi / dx) * np.sin(0.5 * np.pi * s)
    pre = (0.25 / dx) ** 3
    out = np.empty_like(s)
    out[0] = pre * dphi[0] * phi[1] * phi[2]
    out[1] = pre * phi[0] * dphi[1] * phi[2]
    out[2] = pre * phi[0] * phi[1] * dphi[2]
    return out


def _block(idx_i, width=ib.WIDTH):
    ix, iy, iz = idx_i
    return (
        slice(iz - width + 1, iz + width + 1),
        slice(iy - width + 1, iy + width + 1),
        slice(ix - width + 1, ix + width + 1),
    )


def weight_gradient_contraction(eul_field, lag_coef, dw, idx):
    """``out[a, n] = sum_c lag_coef[c, n] sum_block eul_field[c][block] dw[a, ..., n]``."""
    n = idx.shape[1]
    out = np.zeros((3, n))
    for i in range(n):
        blk = _block(idx[:, i])
        for c in range(3):
            e = eul_field[c][blk]
            for a in range(3):
                out[a, i] += lag_coef[c, i] * np.sum(e * dw[a, ..., i])
    return out


# ----------------------------------------------------------------- heave kinematics


class HeaveKinematics:
    """Rigid heave in y: ``v(t) = sum_{m=1}^{M} a_m cos(2 pi m t/T) + b_m sin(2 pi m t/T)``,
    ``Y(t) = int_0^t v``; theta = (a_1..a_M, b_1..b_M). Nodes ``X0 + (0, Y, 0)``, velocity
    ``(0, v, 0)``. Linear in theta."""

    def __init__(self, X0, period, n_modes=8):
        self.X0 = np.asarray(X0, dtype=np.float64)
        self.period = float(period)
        self.n_modes = int(n_modes)

    def basis(self, t):
        m = np.arange(1, self.n_modes + 1)
        om = 2 * np.pi * m / self.period
        cv = np.concatenate([np.cos(om * t), np.sin(om * t)])
        cy = np.concatenate([np.sin(om * t) / om, (1 - np.cos(om * t)) / om])
        return cv, cy

    def state(self, theta, t):
        cv, cy = self.basis(t)
        X = self.X0.copy()
        X[1] += cy @ theta
        V = np.zeros_like(X)
        V[1] = cv @ theta
        return X, V


# ----------------------------------------------------------------- rollout and its adjoint


class Rollout:
    """Forward rollout of the coupled step and its reverse-mode gradient (oracle, FP64)."""

    def __init__(self, grid_size, x_range, nu, omega0, free_stream, kin, k, c, real_t=np.float64,
                 cfl=0.1, width=2, rho=1.0):
        self.grid_size = tuple(grid_size)
        self.x_range = x_range
        self.nu = nu
        self.omega0 = np.asarray(omega0, dtype=real_t)
        self.U = np.asarray(free_stream, dtype=np.float64)
        self.kin = kin
        self.k = k
        self.c = c
        self.real_t = real_t
        self.cfl = cfl
        self.width = width
        self.rho = rho

    def _new_flow(self):
        f = flow_mod.UnboundedNavierStokesFlowSimulator3D(
            self.grid_size, self.x_range, self.nu, cfl=self.cfl, real_t=self.real_t,
            with_forcing=True, with_free_stream_flow=True, flow_density=self.rho,
            penalty_zone_width=self.width,
        )
        f.vorticity_field[...] = self.omega0
        f.recover_velocity(self.U)
        return f

    def forward(self, theta, nsteps, dts=None, record=False):
        """Body forces (nsteps, 3) and the dt of each step; with ``record`` also every
        step's inputs for the reverse sweep."""
        flow = self._new_flow()
        n = self.kin.X0.shape[1]
        vbf = ib.VirtualBoundaryForcing(self.k, self.c, flow.dx, n, self.real_t,
                                        enable_eul_grid_forcing_reset=False)
        forces = np.zeros((nsteps, 3))
        used = np.zeros(nsteps)
        tape = []
        t = 0.0
        for s in range(nsteps):
            dt = dts[s] if dts is not None else flow.compute_stable_timestep()
            used[s] = dt
            omega_n = flow.vorticity_field.copy() if record else None
            u_n = flow.velocity_field.copy() if record else None
            vbf.time_step(dt)
            X, V = self.kin.state(theta, t)
            vbf.compute_interaction_forcing(flow.eul_grid_forcing_field, flow.velocity_field, X, V)
            forces[s] = -vbf.lag_grid_forcing_field.sum(axis=1)
            if record:
                tape.append(dict(omega=omega_n, u=u_n, F=flow.eul_grid_forcing_field.copy(),
                                 X=X, Fl=vbf.lag_grid_forcing_field.copy(),
                                 idx=vbf.nearest_eul_grid_index_to_lag_grid.copy(),
                                 w=vbf.interp_weights.copy(), dt=dt, t=t))
            flow.time_step(dt, self.U)
            t += dt
        self.flow = flow
        self.tape = tape
        return forces, used

    def gradient(self, theta, nsteps, dts, fbar, include_positions=True):
        """d(sum_n fbar_n . f_n)/d theta with dt frozen to ``dts``. ``include_positions=False``
        drops the node-position (dw/dX) part, to show that it matters."""
        self.forward(theta, nsteps, dts=dts, record=True)
        dx = self.flow.dx
        mats = penalise_matrices(self.flow)
        n = self.kin.X0.shape[1]
        wbar = np.zeros((3, *self.grid_size))
        ubar = np.zeros_like(wbar)
        dXbar = np.zeros((3, n))
        dVbar = np.zeros((3, n))
        theta_bar = np.zeros_like(np.asarray(theta, dtype=np.float64))
        shift = dx / 2
        for s in reversed(range(nsteps)):
            rec = self.tape[s]
            wbar, u_n_bar, Fbar = flow_step_adjoint(self.flow, rec["dt"], rec["omega"], rec["u"],
                                                    rec["F"], wbar, ubar, mats)
            Xbar, Vbar, dXbar, dVbar = vbf_adjoint(
                Fbar, rec["u"], rec["X"], rec["Fl"], rec["idx"], rec["w"], fbar[s], dXbar, dVbar,
                self.k, self.c, rec["dt"], dx, shift, u_n_bar)
            ubar = u_n_bar
            cv, cy = self.kin.basis(rec["t"])
            theta_bar += (Xbar[1].sum() if include_positions else 0.0) * cy + Vbar[1].sum() * cv
        return theta_bar


def step_prefactors(flow, dt):
    """pF, pB, pD, pC of the step as the reference forms them
    (``navier_stokes_flow_simulators.py:463,469,482,495``)."""
    rt, dx = flow.real_t, flow.dx
    return (rt(dt / (2 * dx * flow.flow_density)), rt(dt / (2 * dx)),
            rt(flow.kinematic_viscosity * dt / dx / dx), rt(0.5 / dx))


def flow_step_adjoint(flow, dt, omega_n, u_n, F_n, wbar, ubar, mats=None):
    """Reverse of one flow step of the oracle simulator ``flow`` with inputs (omega_n, u_n, F_n):
    (omega_bar after, u_bar after) -> (omega_bar before, u_bar before through the advection
    term, F_bar)."""
    if mats is None:
        mats = penalise_matrices(flow)
    pF, pB, pD, pC = step_prefactors(flow, dt)
    psibar = curl_transpose(ubar, pC)
    gpsi = np.zeros_like(psibar)
    flow.poisson.vector_field_solve(gpsi, psibar)
    g = penalise_apply(wbar + gpsi, mats, transpose=True)
    g = diffusion_transpose(g, pD)
    Bbar = curl_transpose(g, pB)
    w1 = omega_n.copy()
    stencils.update_vorticity_from_velocity_forcing(w1, F_n, pF)
    u_n_bar = cross(w1, Bbar)
    g = g + cross(Bbar, u_n)
    Fbar = curl_transpose(g, pF)
    return g, u_n_bar, Fbar


def vbf_adjoint(Fbar, u, X, Fl, idx, w, fbar_s, dXbar, dVbar, k, c, dt, dx, shift, ubar):
    """Reverse of one virtual-boundary forcing (``VirtualBoundaryForcing.py:149-230``, spreading
    ``EulerianLagrangianGridCommunicator3D.py:318-380``) given the forcing adjoint ``Fbar``:
    returns (X_bar, V_bar, dX_bar, dV_bar of the previous step) and adds the interpolation
    transpose ``dx^3 S(U_bar)`` into ``ubar``."""
    Flbar = ib.interpolate(Fbar, w, idx, dx) / dx**3 - np.asarray(fbar_s).reshape(3, 1)
    dw = cosine_weight_gradient(idx, X, dx, shift)
    Xbar = weight_gradient_contraction(Fbar, Fl, dw, idx)
    dXbar = dXbar + k * Flbar
    Ubar = dVbar + c * Flbar
    Vbar = -Ubar
    ib.spread(ubar, Ubar * dx**3, w, idx)
    Xbar = Xbar + dx**3 * weight_gradient_contraction(u, Ubar, dw, idx)
    return Xbar, Vbar, dXbar, dt * dXbar


def time_average_weights(dts):
    """Weights of the time-averaged body force ``sum_n f_n dt_n / sum_n dt_n``."""
    dts = np.asarray(dts, dtype=np.float64)
    return dts / dts.sum()
