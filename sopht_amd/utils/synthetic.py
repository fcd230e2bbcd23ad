"""Seeded synthetic inputs for the flow-step workloads (host numpy, float64).

The reference ships no data generator; these define the benchmark and parity inputs:

* ``random_phase_vorticity``: ``omega = curl A`` with
  ``A = sum_{0<|k|<=K} a_k |k|^-2 cos(2 pi k.x / L + phi_k)``, ``a_k ~ N(0,1)^3``,
  ``phi_k ~ U[0, 2 pi)``; the curl is analytic,
  ``omega = sum -(2 pi / L) sin(2 pi k.x / L + phi_k) (k x a_k) |k|^-2``, rescaled so that
  ``max_cells |omega| = max_vorticity``. Evaluated separably (three small contractions)
  so a 256^3 field costs seconds.
* ``fibonacci_sphere`` / ``fibonacci_ellipsoid``: Fibonacci lattice
  ``z_i = 1 - (2i+1)/n``, ``phi_i = i pi (3 - sqrt 5)``, scaled per axis.
"""

from __future__ import annotations

import numpy as np


def wavevectors(max_wavenumber: int) -> np.ndarray:
    """Integer wavevectors ``0 < |k| <= K`` in lexicographic (kz, ky, kx) order, rows (kx, ky, kz)."""
    r = np.arange(-max_wavenumber, max_wavenumber + 1)
    kz, ky, kx = np.meshgrid(r, r, r, indexing="ij")
    k = np.stack([kx.ravel(), ky.ravel(), kz.ravel()], axis=1)
    n2 = (k**2).sum(1)
    return k[(n2 > 0) & (n2 <= max_wavenumber**2)]


def random_phase_vorticity(
    grid_size, x_range=1.0, max_wavenumber=6, seed=42, max_vorticity=10.0
) -> np.ndarray:
    """Seeded divergence-free vorticity ``(3, Nz, Ny, Nx)`` float64 on cell centres."""
    nz, ny, nx = grid_size
    dx = x_range / nx
    rng = np.random.default_rng(seed)
    k = wavevectors(max_wavenumber)
    a = rng.standard_normal((k.shape[0], 3))
    phi = rng.uniform(0.0, 2.0 * np.pi, k.shape[0])
    kk = (k**2).sum(1).astype(np.float64)
    # coefficient of exp(i theta) in  -(2pi/L) sin(theta) (k x a)/|k|^2 :  sin = Im(e^{i theta})
    kxa = np.cross(k.astype(np.float64), a) / kk[:, None]
    coef = -(2.0 * np.pi / x_range) * kxa * np.exp(1j * phi)[:, None]  # (M, 3)
    K = max_wavenumber
    m = 2 * K + 1
    out = np.empty((3, nz, ny, nx), dtype=np.float64)
    xs = (np.arange(nx) + 0.5) * dx
    ys = (np.arange(ny) + 0.5) * dx
    zs = (np.arange(nz) + 0.5) * dx
    r = np.arange(-K, K + 1)
    ex = np.exp(2j * np.pi * np.outer(xs, r) / x_range)  # (nx, m)
    ey = np.exp(2j * np.pi * np.outer(ys, r) / x_range)
    ez = np.exp(2j * np.pi * np.outer(zs, r) / x_range)
    for c in range(3):
        C = np.zeros((m, m, m), dtype=np.complex128)  # [kz, ky, kx]
        C[k[:, 2] + K, k[:, 1] + K, k[:, 0] + K] = coef[:, c]
        t = np.tensordot(C, ex, axes=([2], [1]))  # (kz, ky, x)
        t = np.tensordot(t, ey, axes=([1], [1]))  # (kz, x, y)
        t = np.tensordot(t, ez, axes=([0], [1]))  # (x, y, z)
        out[c] = np.transpose(t.imag, (2, 1, 0))
    mag = np.sqrt((out**2).sum(0)).max()
    out *= max_vorticity / mag
    return out


def fibonacci_unit_sphere(num_points: int) -> np.ndarray:
    """``(3, n)`` Fibonacci-lattice points on the unit sphere."""
    i = np.arange(num_points, dtype=np.float64)
    z = 1.0 - (2.0 * i + 1.0) / num_points
    phi = i * np.pi * (3.0 - np.sqrt(5.0))
    rxy = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([rxy * np.cos(phi), rxy * np.sin(phi), z])


def fibonacci_sphere(num_points, radius, center) -> np.ndarray:
    return np.asarray(center, dtype=np.float64).reshape(3, 1) + radius * fibonacci_unit_sphere(num_points)


def fibonacci_ellipsoid(num_points, semi_axes, center) -> np.ndarray:
    s = np.asarray(semi_axes, dtype=np.float64).reshape(3, 1)
    return np.asarray(center, dtype=np.float64).reshape(3, 1) + s * fibonacci_unit_sphere(num_points)


class HeavingRod:
    """Prescribed heave of a straight rod (the flexible-filament workload's base motion,
    applied to the whole rod): every node at ``rest + (0, A sin(2 pi f t), 0)`` with velocity
    ``(0, 2 pi f A cos(2 pi f t), 0)``; directors and angular velocities stay as built.
    Stands in for the rod's own time stepper on the host, which is outside this package."""

    def __init__(self, rod, amplitude=0.05, frequency=1.0):
        self.rod = rod
        self.amplitude = float(amplitude)
        self.frequency = float(frequency)
        self.rest = rod.position_collection.copy()

    def set_time(self, t: float) -> None:
        w = 2.0 * np.pi * self.frequency
        self.rod.position_collection[...] = self.rest
        self.rod.position_collection[1] += self.amplitude * np.sin(w * t)
        self.rod.velocity_collection[...] = 0.0
        self.rod.velocity_collection[1] = self.amplitude * w * np.cos(w * t)


def rod_coupled_step(flow, interactor, flow_forces, rod, kinematics, rod_dt, free_stream_velocity, dt_prefac=0.25):
    """One flow step of the flexible-filament driver (``flow_past_rod_case.py:231-250``):
    the flow dt (CFL with ``dt_prefac``), ``int(dt / min(dt, rod_dt))`` rod substeps, each
    moving the rod (here: its prescribed kinematics), applying the flow forces to it (the
    interactor's forces and torques, without spreading) and advancing the interactor, then
    one interactor call (spreading) and the flow step. Returns the flow dt."""
    dt = flow.compute_stable_timestep(dt_prefac=dt_prefac)
    substeps = int(dt / min(dt, rod_dt))
    local_dt = dt / substeps
    t = flow.time
    for _ in range(substeps):
        t = t + local_dt
        kinematics.set_time(t)
        rod.external_forces[...] = 0.0
        rod.external_torques[...] = 0.0
        flow_forces.apply_forces(rod, t)
        interactor.time_step(dt=local_dt)
    interactor()
    flow.time_step(dt=dt, free_stream_velocity=free_stream_velocity)
    return dt
