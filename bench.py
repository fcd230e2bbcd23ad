#!/usr/bin/env python3
"""Throughput of the coupled 3D IB flow step on one MI355X.

Workload (the plain method on the GPU at 256^3 in FP32): x_range 1, nu = 5e-4, free stream
(1, 0, 0) every step, flow density 1, penalty width 2; initial vorticity = curl of a seeded
random-phase vector potential (|k| <= 8, amplitude ~ k^-2, seed 42) scaled to
max|omega| = 10, then one velocity recovery; a fixed sphere of radius 0.08 at the centre
with 3000 Fibonacci-lattice forcing points, virtual-boundary stiffness -5e4 and damping
-10 (the reference's sign convention), both rescaled by the Lagrangian spacing squared
as ImmersedBodyFlowInteraction does.

One step is the reference driver's loop body: ``dt = flow_sim.compute_stable_timestep()``
(CFL 0.1), ``interactor.time_step(dt)``, ``interactor()``, ``flow_sim.time_step(dt, U)``.
Inputs are resident on the device before timing starts; W untimed warm-up steps, then K
steps between two device synchronisations.

Prints one JSON line with the throughput in cell-updates/s, the per-kernel roofline of
the dominant kernel (its average duration from HIP events recorded on its own stream
during the timed steps) and a CPU baseline: the numpy/scipy restatement of the same loop
timed on this host on a bounded sample of the same workload.

Secondary objects on the same line: "batched" (64 small coupled instances in one batched
step against the same instances stepped one after another), "adjoint" (forward rollout
against forward with checkpoints plus the adjoint sweep, 128^3, 200 steps) and
"fast_diagonalisation" (the Neumann fast-diagonalisation vector solve at 256^3 FP32 against
the FP32 MFMA peak, the flow step with that solver, and the numpy oracle of the solve),
"rod_coupling" (the flexible-filament driver loop at 128^3 with a heaving rod: rod substeps with
the flow forces, one spreading per flow step), "other_sizes" (the sphere workload at 192^3
fp32 and fp64 and at 256^3 fp64, each with its roofline), "filtered_step" (the 256^3 step
with the order-5 convolution vorticity filter of the flexible-filament example) and
"standalone_factories" (the drop-in stencil factories off the step path at 256^3 FP32: ENO3
flux and its adjoint, vorticity stretching, curl, diffusion, each against HBM peak).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
WORKLOAD = "coupled IB flow step, 256^3 fp32, sphere r=0.08 (3000 Fibonacci points), nu=5e-4, U=(1,0,0), CFL 0.1"
# frozen algorithmic traffic of the whole step in words per cell: the minimal reads and
# writes of the step's fields (vorticity, velocity, forcing, stream function) per cell
STEP_WORDS_PER_CELL = 76


def stage_words(nz, ny, nx):
    """Algorithmic words moved per launch of each stage of sopht_ns3d_step, per cell
    (reads + writes of the arrays each kernel must touch, each counted once)."""
    c = (nx + 1) / nx  # complex columns kx = 0..nx per real row of nx
    ghat = (nz + 1) * (ny + 1) * (nx + 1) / (nz * ny * nx)
    return {
        "vorticity_forcing_advect": 9 + 3,     # read w, u, F; write w2 (into psi)
        "vorticity_diffuse_penalise": 3 + 6,   # read w2; write w, F = 0
        "poisson_x_r2c": 3 * (1 + 2 * c),      # real rows in, half spectra out
        "poisson_y_fwd": 3 * (2 * c + 4 * c),  # y < ny in, all 2ny out
        "poisson_z_conv": 3 * (4 * c + 4 * c) + ghat,  # z < nz in and out, G^ octant
        "poisson_y_inv": 3 * (4 * c + 2 * c),
        "poisson_x_c2r": 3 * (2 * c + 1),
        # three-pass schedule (nx = ny): x rows with the inner y sums, and their inverse
        "poisson_xy_fwd": 3 * (1 + 4 * c),     # real rows in, all 2ny rows of the slab out
        "poisson_yx_inv": 3 * (4 * c + 1),
        "velocity_curl": 3 + 3,                # read psi; write u
    }


def hbm_traffic(stage):
    """HBM bytes per launch of a stage at 256^3 FP32, from the committed rocprofv3 summary
    (FETCH_SIZE and WRITE_SIZE runs, calibrated with the copy probe); None if absent."""
    path = ROOT / "profiles" / "flow_step_256_fp32_hbm_traffic.md"
    try:
        for line in path.read_text().splitlines():
            cols = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cols) == 7 and cols[0] == stage:
                return int(cols[4])
    except (OSError, ValueError):
        return None
    return None


def build(n, real_t, seed=42, n_lag=3000, radius=0.08, nu=5e-4, K=8, flow_kw=None):
    from sopht_amd import simulator as sps
    from sopht_amd.utils import synthetic

    import torch

    flow = sps.UnboundedNavierStokesFlowSimulator3D(
        grid_size=(n, n, n), x_range=1.0, kinematic_viscosity=nu, cfl=0.1, real_t=real_t,
        with_forcing=True, with_free_stream_flow=True, **(flow_kw or {}),
    )
    w0 = synthetic.random_phase_vorticity((n, n, n), 1.0, K, seed, 10.0)
    flow.vorticity_field.copy_(torch.from_numpy(w0.astype(real_t)))
    del w0
    U = np.array([1.0, 0.0, 0.0])
    flow.recover_velocity(U)
    body = sps.RigidBody((0.5, 0.5, 0.5), radius)
    inter = sps.RigidBodyFlowInteraction(
        body, flow.eul_grid_forcing_field, flow.velocity_field, -5e4, -10.0, flow.dx, 3,
        sps.FibonacciSphereForcingGrid, real_t=real_t, num_lag_nodes=n_lag,
    )
    return flow, inter, U


def step(flow, inter, U):
    dt = flow.compute_stable_timestep()
    inter.time_step(dt)
    inter()
    flow.time_step(dt, free_stream_velocity=U)
    return dt


def cpu_baseline(n, real_t, steps, n_lag=3000):
    """The oracle (numpy + scipy.fft with all host cores for the FFTs) on the same loop."""
    from oracle import flow as OF
    from oracle import poisson as OP
    from sopht_amd import simulator as sps
    from sopht_amd.utils import synthetic

    cores = len(os.sched_getaffinity(0))
    OP.WORKERS = cores
    o = OF.UnboundedNavierStokesFlowSimulator3D(
        (n, n, n), 1.0, 5e-4, cfl=0.1, real_t=real_t, with_forcing=True, with_free_stream_flow=True
    )
    o.vorticity_field[...] = synthetic.random_phase_vorticity((n, n, n), 1.0, 8, 42, 10.0).astype(real_t)
    U = np.array([1.0, 0.0, 0.0])
    o.recover_velocity(U)
    body = sps.RigidBody((0.5, 0.5, 0.5), 0.08)
    oi = OF.ImmersedBodyFlowInteraction(o, sps.FibonacciSphereForcingGrid(3, body, n_lag), -5e4, -10.0, real_t)
    t0 = time.perf_counter()
    for _ in range(steps):
        dt = o.compute_stable_timestep()
        oi.time_step(dt)
        oi()
        o.time_step(dt, free_stream_velocity=U)
    el = time.perf_counter() - t0
    return {
        "value": n**3 * steps / el,
        "unit": "cell-updates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{steps} steps of the same {n}^3 loop (numpy stencils single-threaded, scipy.fft with {cores} workers), {el:.1f} s",
    }


# the batched workload: 64 uneven members (sizes and counts)
BATCH_SIZES = [(48, 16), (64, 24), (96, 16), (128, 8)]


def batched_members(seed=7, n_lag=1000, nu=1e-3, K=6):
    """64 independent FP32 members: per-member seed (random-phase vorticity |k| <= 6,
    amplitude ~ k^-2, max|omega| = 10), a fixed sphere of radius ~ U[0.05, 0.12] at the
    centre with 1000 Fibonacci points, free stream (U, 0, 0) with U ~ U[0.5, 1.5]."""
    import torch

    from sopht_amd import simulator as sps
    from sopht_amd.utils import synthetic

    rng = np.random.default_rng(seed)
    members = []
    k = 0
    for n, count in BATCH_SIZES:
        for _ in range(count):
            flow = sps.UnboundedNavierStokesFlowSimulator3D(
                grid_size=(n, n, n), x_range=1.0, kinematic_viscosity=nu, cfl=0.1, real_t=np.float32,
                with_forcing=True, with_free_stream_flow=True,
            )
            w0 = synthetic.random_phase_vorticity((n, n, n), 1.0, K, 1000 + k, 10.0)
            flow.vorticity_field.copy_(torch.from_numpy(w0.astype(np.float32)))
            U = np.array([rng.uniform(0.5, 1.5), 0.0, 0.0])
            flow.recover_velocity(U)
            body = sps.RigidBody((0.5, 0.5, 0.5), rng.uniform(0.05, 0.12))
            inter = sps.RigidBodyFlowInteraction(
                body, flow.eul_grid_forcing_field, flow.velocity_field, -5e4, -10.0, flow.dx, 3,
                sps.FibonacciSphereForcingGrid, real_t=np.float32, num_lag_nodes=n_lag,
            )
            members.append((flow, inter, U))
            k += 1
    return members


def batched_vs_sequential(steps, seq_steps, warmup=2):
    """Cell-updates/s of the 64 members advanced by the batched step (one hipGraph replay
    per step) and by stepping them one after another with the reference driver loop."""
    import logging

    import torch

    from sopht_amd.simulator import BatchedFlowSimulation

    logging.getLogger("sopht_amd").setLevel(logging.ERROR)
    members = batched_members()
    cells = sum(int(np.prod(f.grid_size)) for f, _, _ in members)
    for _ in range(warmup):
        for flow, inter, U in members:
            step(flow, inter, U)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(seq_steps):
        for flow, inter, U in members:
            step(flow, inter, U)
    torch.cuda.synchronize()
    seq = time.perf_counter() - t0
    sim = BatchedFlowSimulation(members)
    schedules = sim.poisson_schedules()
    sim.step(1 + warmup)
    sim.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.step(steps)
    sim.synchronize()
    bat = time.perf_counter() - t0
    sim.close()
    seq_rate = cells * seq_steps / seq
    bat_rate = cells * steps / bat
    return {
        "members": len(members),
        "sizes": [f"{count}x{n}^3" for n, count in BATCH_SIZES],
        "cells_per_step": cells,
        "batched_cell_updates_per_s": bat_rate,
        "batched_ms_per_step": round(1e3 * bat / steps, 3),
        "sequential_cell_updates_per_s": seq_rate,
        "sequential_ms_per_step": round(1e3 * seq / seq_steps, 3),
        "speedup": round(bat_rate / seq_rate, 2),
        "steps": steps,
        "sequential_steps": seq_steps,
        "poisson_schedules": schedules,
    }


ADJ_WORKLOAD = ("heaving ellipsoid (0.1, 0.05, 0.05), 2000 Fibonacci points, 128^3 fp32, nu=1e-3, "
                "U=(1,0,0), 200 steps, 16 Fourier heave coefficients ~ N(0, 0.1), checkpoint every 20")


def adjoint_vs_forward(n=128, steps=200, checkpoint=20, seed=5):
    """Time of the forward rollout (the driver loop with CFL dt) against the forward pass with
    checkpoints plus the adjoint (recompute + reverse sweep), for the time-averaged body
    x-force of a heaving ellipsoid; both from the same start state."""
    import logging

    import torch

    from sopht_amd import simulator as sps
    from sopht_amd.numeric.immersed_boundary_ops import VirtualBoundaryForcing
    from sopht_amd.utils import synthetic

    logging.getLogger("sopht_amd").setLevel(logging.ERROR)
    real_t = np.float32
    n_lag, axes = 2000, (0.1, 0.05, 0.05)
    flow = sps.UnboundedNavierStokesFlowSimulator3D(
        grid_size=(n, n, n), x_range=1.0, kinematic_viscosity=1e-3, cfl=0.1, real_t=real_t,
        with_forcing=True, with_free_stream_flow=True)
    U = np.array([1.0, 0.0, 0.0])
    flow.recover_velocity(U)
    grid = sps.FibonacciEllipsoidForcingGrid(3, sps.RigidBody((0.5, 0.5, 0.5)), n_lag, axes)
    ds = grid.get_maximum_lagrangian_grid_spacing()
    inter = VirtualBoundaryForcing(-5e4 * ds**2, -10.0 * ds**2, 3, flow.dx, n_lag, real_t,
                                   enable_eul_grid_forcing_reset=False)
    kin = sps.HeaveKinematics(synthetic.fibonacci_ellipsoid(n_lag, axes, (0.5, 0.5, 0.5)), period=0.25)
    ro = sps.CoupledFlowRollout(flow, inter, kin, U, steps, checkpoint_every=checkpoint)
    theta = torch.tensor(np.random.default_rng(seed).normal(0.0, 0.1, 16), device="cuda",
                         requires_grad=True)
    # warm-up (allocations, plans, the shadow interactor)
    wro = sps.CoupledFlowRollout(flow, inter, kin, U, 2, checkpoint_every=1)
    sps.time_averaged_force(sps.body_forces(theta, wro), wro.dts).backward()
    del wro
    theta.grad = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f_plain = ro.run(theta, record=False)
    torch.cuda.synchronize()
    fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    forces = sps.body_forces(theta, ro)
    J = sps.time_averaged_force(forces, ro.dts)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    J.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    same = bool(torch.equal(f_plain, forces.detach()))
    g = theta.grad.detach().cpu().numpy()
    return {
        "workload": ADJ_WORKLOAD,
        "forward_s": round(fwd, 4),
        "forward_recorded_s": round(t1 - t0, 4),
        "adjoint_s": round(t2 - t1, 4),
        "forward_plus_adjoint_s": round(t2 - t0, 4),
        "ratio_fwd_adj_to_fwd": round((t2 - t0) / fwd, 3),
        "forward_ms_per_step": round(1e3 * fwd / steps, 3),
        "objective": float(J.detach()),
        "grad_norm": float(np.linalg.norm(g)),
        "forward_recorded_bitwise_equal": same,
        "gradient_check": "tests/test_gpu_adjoint.py: 64^3 fp64 gradient vs central differences (h=1e-4)",
    }


ROD_WORKLOAD = ("heaving flexible-filament coupling, 128^3 fp32, nu=1e-3, U=(1,0,0): rod length 0.3, 64 "
                "elements, radius 0.005, base (0.3, 0.5, 0.5) along x, heave 0.05 sin(2 pi t) in y, forcing at "
                "the 65 nodes, rod substeps of 0.01 l_elem with the flow forces applied, CFL dt x 0.25")


def rod_coupled(n=128, steps=20, warmup=2):
    """The flexible-filament driver loop (flow_past_rod_case.py:231-250) on a prescribed heave:
    per flow step, rod substeps (kinematics on the host, the flow forces on the rod: Lagrangian
    forcing on the device, host transfer), one spreading, the flow step. Reported against the
    flow step alone at the same size."""
    import logging

    import torch

    from sopht_amd import simulator as sps
    from sopht_amd.utils import synthetic

    logging.getLogger("sopht_amd").setLevel(logging.ERROR)
    real_t = np.float32
    flow = sps.UnboundedNavierStokesFlowSimulator3D(
        grid_size=(n, n, n), x_range=1.0, kinematic_viscosity=1e-3, cfl=0.1, real_t=real_t,
        with_forcing=True, with_free_stream_flow=True)
    w0 = synthetic.random_phase_vorticity((n, n, n), 1.0, 6, 42, 10.0)
    flow.vorticity_field.copy_(torch.from_numpy(w0.astype(real_t)))
    U = np.array([1.0, 0.0, 0.0])
    flow.recover_velocity(U)
    n_elem, length = 64, 0.3
    rod = sps.StraightRod(n_elem, np.array([0.3, 0.5, 0.5]), np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0]),
                          length, 0.005, density=1.0)
    inter = sps.CosseratRodFlowInteraction(rod, flow.eul_grid_forcing_field, flow.velocity_field, -2e4, -1e2, flow.dx,
                                           3, sps.CosseratRodNodalForcingGrid, real_t=real_t)
    kin = synthetic.HeavingRod(rod, amplitude=0.05, frequency=1.0)
    ff = sps.FlowForces(inter)
    rod_dt = 0.01 * length / n_elem
    for _ in range(warmup):
        synthetic.rod_coupled_step(flow, inter, ff, rod, kin, rod_dt, U)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    subs = 0
    for _ in range(steps):
        dt = synthetic.rod_coupled_step(flow, inter, ff, rod, kin, rod_dt, U)
        subs += int(dt / min(dt, rod_dt))
    torch.cuda.synchronize()
    coupled = time.perf_counter() - t0
    inter.check_support_in_domain()
    t0 = time.perf_counter()
    for _ in range(steps):
        flow.time_step(dt=1e-4, free_stream_velocity=U)
    torch.cuda.synchronize()
    plain = time.perf_counter() - t0
    return {
        "workload": ROD_WORKLOAD,
        "ms_per_flow_step": round(1e3 * coupled / steps, 4),
        "cell_updates_per_s": n**3 * steps / coupled,
        "rod_substeps_per_flow_step": round(subs / steps, 2),
        "ms_per_rod_substep": round(1e3 * (coupled - plain) / max(subs, 1), 4),
        "flow_step_alone_ms": round(1e3 * plain / steps, 4),
        "steps": steps,
        "note": "rod mechanics prescribed on the host (pyelastica absent); parity of the rod dynamics unpinned",
    }


def plain_rate(n, real_t, steps=20, warmup=3, flow_kw=None):
    """The coupled sphere loop at another size or precision (the 192^3 / 256^3 workload
    family), or with other simulator options (flow_kw): ms per step, cell-updates/s and the
    step's fraction of the 76-word HBM model, with the dominant kernel's roofline."""
    import ctypes as C

    import torch

    from sopht_amd import _lib as L

    flow, inter, U = build(n, real_t, flow_kw=flow_kw)
    for _ in range(warmup):
        step(flow, inter, U)
    lib = L.lib()
    L.check(lib.sopht_ns3d_enable_timing(flow._plan, 1))
    ms = (C.c_double * 16)()
    cnt = C.c_int64()
    lib.sopht_ns3d_stage_times(flow._plan, ms, 16, C.byref(cnt), 1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step(flow, inter, U)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    nst = lib.sopht_ns3d_stage_times(flow._plan, ms, 16, C.byref(cnt), 0)
    names = [lib.sopht_ns3d_stage_name(flow._plan, i).decode() for i in range(nst)]
    avg = {names[i]: ms[i] / max(cnt.value, 1) for i in range(nst)}
    word = np.dtype(real_t).itemsize
    words = stage_words(n, n, n)
    if flow_kw and flow_kw.get("filter_vorticity"):
        # the filter runs inside the second stage: order-k convolution, one read and one write
        # of the three components per axis (penalisation planes neglected)
        words["vorticity_diffuse_penalise"] += 3 * 6
    dom = max(avg, key=avg.get)
    gbs = words[dom] * n**3 * word / (avg[dom] * 1e-3) / 1e9
    ms_step = 1e3 * el / steps
    step_gbs = STEP_WORDS_PER_CELL * word * n**3 / (ms_step * 1e-3) / 1e9
    return {
        "grid": n, "dtype": "f32" if real_t == np.float32 else "f64",
        "ms_per_step": round(ms_step, 4), "cell_updates_per_s": n**3 * steps / el,
        "step_hbm_frac_76w": round(step_gbs / HBM_PEAK_GBS, 4),
        "stage_ms": {k: round(v, 4) for k, v in avg.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)},
        "steps": steps,
    }


FD_WORKLOAD = "Neumann fast-diagonalisation vector Poisson solve, 256^3 fp32 (3 components)"
FP32_MFMA_PEAK_TFLOPS = 157.3


def fastdiag_solve(n=256, reps=20, cpu=True):
    """The fast-diagonalisation (Neumann) vector solve: six dense per-axis transforms of
    2 n^4 flops per component each (rocBLAS GEMMs on the MFMA units) and the spectral scaling,
    timed with HIP events on the launch stream; the 256^3 flow step with this solver; and the
    oracle's numpy solve of the same field on the host as the CPU baseline."""
    import torch

    from sopht_amd import simulator as sps
    from sopht_amd.numeric.eulerian_grid_ops import FastDiagPoissonSolver3D

    s = FastDiagPoissonSolver3D(n, n, n, dx=1.0 / n, real_t=np.float32)
    gen = torch.Generator(device="cuda").manual_seed(11)
    rhs = torch.randn((3, n, n, n), dtype=torch.float32, device="cuda", generator=gen)
    out = torch.empty_like(rhs)
    for _ in range(3):
        s.vector_field_solve(out, rhs)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(reps):
        s.vector_field_solve(out, rhs)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / reps
    flops = 6 * 3 * 2.0 * n**4
    tflops = flops / (ms * 1e-3) / 1e12
    del s, out
    torch.cuda.empty_cache()
    flow = sps.UnboundedNavierStokesFlowSimulator3D(
        grid_size=(n, n, n), x_range=1.0, kinematic_viscosity=5e-4, real_t=np.float32,
        with_forcing=True, with_free_stream_flow=True, poisson_solver_type="fast_diagonalisation")
    flow.vorticity_field.copy_(rhs)
    U = np.array([1.0, 0.0, 0.0])
    for _ in range(3):
        flow.time_step(1e-4, free_stream_velocity=U)
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(reps):
        flow.time_step(1e-4, free_stream_velocity=U)
    ev[1].record()
    torch.cuda.synchronize()
    step_ms = ev[0].elapsed_time(ev[1]) / reps
    result = {
        "workload": FD_WORKLOAD,
        "ms_per_solve": round(ms, 4),
        "roofline": {
            "bound": "mfma", "achieved": round(tflops, 2), "peak": FP32_MFMA_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(tflops / FP32_MFMA_PEAK_TFLOPS, 4),
            "flops_per_solve": int(flops), "traffic": None,
        },
        "flow_step_ms": round(step_ms, 4),
        "flow_step_cell_updates_per_s": n**3 / (step_ms * 1e-3),
    }
    if cpu:
        from threadpoolctl import threadpool_info

        from oracle import fastdiag as FD

        h = rhs.cpu().numpy()
        o = FD.FastDiagPoissonSolver3D(n, n, n, np.float32(1.0 / n), np.float32)
        ho = np.zeros_like(h)
        t0 = time.perf_counter()
        o.vector_field_solve(ho, h)
        dt = time.perf_counter() - t0
        threads = max((i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"),
                      default=1)
        result["cpu_baseline"] = {
            "value": round(1e3 * dt, 1), "unit": "ms per vector solve", "cores": threads, "kind": "port",
            "sample": f"one {n}^3 fp32 vector solve of the oracle (numpy eig + tensordot, threaded BLAS)",
        }
    return result


FILTER_WORKLOAD = ("coupled IB flow step, 256^3 fp32, with the vorticity filter of the flexible-filament "
                   "example (order 5, convolution)")


def filtered_step(steps):
    """The 256^3 FP32 sphere step with the Laplacian vorticity filter on, order 5
    convolution (flow_past_rod_case.py:103-104): ms per step against the same step without it."""
    kw = {"filter_vorticity": True, "filter_setting_dict": {"order": 5, "type": "convolution"}}
    out = plain_rate(256, np.float32, steps, flow_kw=kw)
    out["workload"] = FILTER_WORKLOAD
    return out


def other_sizes(steps):
    """The sphere workload at 192^3 (fp32, fp64) and 256^3 fp64."""
    import torch

    out = []
    for n, real_t in ((192, np.float32), (192, np.float64), (256, np.float64)):
        out.append(plain_rate(n, real_t, steps))
        torch.cuda.empty_cache()
    return out


def standalone_factories(n=256, reps=20):
    """The drop-in factories off the step path, one (n, n, n) FP32 field each, timed with HIP
    events on the stream they launch on. Algorithmic words per cell: ENO3 flux 6 (field,
    three velocity components, flux read and written), its adjoint 13 (flux adjoint, field,
    velocity, field adjoint and velocity adjoint read and written), vorticity stretching 9,
    curl 6, vector diffusion flux 6."""
    import torch

    from sopht_amd.numeric import eulerian_grid_ops as ops

    rt = np.float32
    gen = torch.Generator(device="cuda").manual_seed(3)
    r = lambda *s: torch.randn(s, dtype=torch.float32, device="cuda", generator=gen)  # noqa: E731
    f, flux, fa = r(n, n, n), r(n, n, n), r(n, n, n)
    u, w, ua, out = r(3, n, n, n), r(3, n, n, n), r(3, n, n, n), r(3, n, n, n)
    eno3 = ops.gen_advection_flux_conservative_eno3_pyst_kernel_3d(rt)
    eno3_adj = ops.gen_advection_flux_conservative_eno3_adjoint_kernel_3d(rt)
    stretch = ops.gen_vorticity_stretching_flux_pyst_kernel_3d(rt)
    curl = ops.gen_curl_pyst_kernel_3d(rt)
    diff = ops.gen_diffusion_flux_pyst_kernel_3d(rt, field_type="vector")
    cases = (
        ("eno3_flux", 6, lambda: eno3(flux, f, u, 0.1)),
        ("eno3_flux_adjoint", 13, lambda: eno3_adj(fa, ua, flux, f, u, 0.1)),
        ("vorticity_stretching_flux", 9, lambda: stretch(out, w, u, 0.1)),
        ("curl", 6, lambda: curl(curl=out, field=w, prefactor=0.1)),
        ("diffusion_flux_vector", 6, lambda: diff(out, w, 0.1)),
    )
    res = {}
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    for name, words, fn in cases:
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        ev[0].record()
        for _ in range(reps):
            fn()
        ev[1].record()
        torch.cuda.synchronize()
        ms = ev[0].elapsed_time(ev[1]) / reps
        gbs = words * 4 * n**3 / (ms * 1e-3) / 1e9
        res[name] = {"ms": round(ms, 4), "words_per_cell": words, "achieved_gbs": round(gbs, 1),
                     "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return {"grid": [n, n, n], "dtype": "f32", "reps": reps, "ops": res}


def copy_probe(n, reps):
    """A plain device copy of known size (one (n, n, n) fp32 field per launch, read once
    and written once) through the library's elementwise copy kernel: the calibration run
    for the HBM byte counters (FETCH_SIZE, WRITE_SIZE) and the achievable copy bandwidth."""
    import torch

    from sopht_amd import _lib as L

    torch.cuda.set_device(0)
    src = torch.rand((n, n, n), device="cuda", dtype=torch.float32)
    dst = torch.empty_like(src)
    st = L.strides3(src)
    args = (L.dtype_tag(np.float32), L.ptr(dst), st, L.ptr(src), st, n, n, n, L.stream_ptr())
    for _ in range(3):
        L.call("sopht_elementwise_copy_3d", *args)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(reps):
        L.call("sopht_elementwise_copy_3d", *args)
    ev[1].record()
    torch.cuda.synchronize()
    if not torch.equal(src, dst):
        print("copy probe mismatch", file=sys.stderr)
        return 1
    ms = ev[0].elapsed_time(ev[1]) / reps
    nbytes = src.numel() * 4
    print(json.dumps({
        "probe": "sopht_elementwise_copy_3d", "grid": [n, n, n], "bytes_read": nbytes,
        "bytes_written": nbytes, "launches": reps + 3, "ms_per_launch": round(ms, 4),
        "copy_gbs": round(2 * nbytes / (ms * 1e-3) / 1e9, 1),
    }))
    return 0


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--precision", choices=["single", "double"], default="single")
    ap.add_argument("--cpu-steps", type=int, default=2, help="steps of the CPU baseline sample (0 = skip)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="REPS",
                    help="only time REPS launches of a plain (grid^3 fp32) copy kernel and exit")
    ap.add_argument("--batch-steps", type=int, default=20,
                    help="steps of the batched 64-member comparison (0 = skip)")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched comparison")
    ap.add_argument("--adjoint-steps", type=int, default=200,
                    help="steps of the forward vs forward+adjoint comparison (0 = skip)")
    ap.add_argument("--adjoint-only", action="store_true", help="run only the adjoint comparison")
    ap.add_argument("--fastdiag-reps", type=int, default=20,
                    help="timed fast-diagonalisation solves and steps (0 = skip)")
    ap.add_argument("--fastdiag-only", action="store_true", help="run only the fast-diagonalisation solve")
    ap.add_argument("--rod-steps", type=int, default=20,
                    help="flow steps of the heaving-filament coupling at 128^3 (0 = skip)")
    ap.add_argument("--rod-only", action="store_true", help="run only the heaving-filament coupling")
    ap.add_argument("--sizes-steps", type=int, default=20,
                    help="steps of the 192^3 fp32/fp64 and 256^3 fp64 lines (0 = skip)")
    ap.add_argument("--sizes-only", action="store_true", help="run only the other sizes and precisions")
    ap.add_argument("--factory-reps", type=int, default=20,
                    help="timed launches of each standalone factory at 256^3 fp32 (0 = skip)")
    ap.add_argument("--filter-steps", type=int, default=20,
                    help="steps of the 256^3 step with the order-5 convolution filter (0 = skip)")
    ap.add_argument("--filter-only", action="store_true", help="run only the filtered step")
    ap.add_argument("--factory-only", action="store_true", help="run only the standalone factories")
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported "
            "(multi-GPU sharding is designed, not built).",
            file=sys.stderr,
        )
        return 2
    if args.copy_probe > 0:
        return copy_probe(args.grid, args.copy_probe)
    if args.adjoint_only:
        import torch

        torch.cuda.set_device(0)
        print(json.dumps(adjoint_vs_forward(steps=args.adjoint_steps)))
        return 0
    if args.fastdiag_only:
        import torch

        torch.cuda.set_device(0)
        print(json.dumps(fastdiag_solve(reps=max(1, args.fastdiag_reps), cpu=args.cpu_steps > 0)))
        return 0
    if args.rod_only:
        import torch

        torch.cuda.set_device(0)
        print(json.dumps(rod_coupled(steps=max(1, args.rod_steps))))
        return 0
    if args.sizes_only:
        import torch

        torch.cuda.set_device(0)
        print(json.dumps(other_sizes(max(1, args.sizes_steps))))
        return 0
    if args.filter_only:
        import torch

        torch.cuda.set_device(0)
        print(json.dumps(filtered_step(max(1, args.filter_steps))))
        return 0
    if args.factory_only:
        import torch

        torch.cuda.set_device(0)
        print(json.dumps(standalone_factories(reps=max(1, args.factory_reps))))
        return 0
    if args.batch_only:
        import torch

        torch.cuda.set_device(0)
        print(json.dumps(batched_vs_sequential(args.batch_steps, max(2, args.batch_steps // 4))))
        return 0

    import torch

    from sopht_amd import _lib as L

    torch.cuda.set_device(0)
    real_t = np.float32 if args.precision == "single" else np.float64
    n = args.grid
    flow, inter, U = build(n, real_t)
    for _ in range(args.warmup):
        step(flow, inter, U)
    lib = L.lib()
    L.check(lib.sopht_ns3d_enable_timing(flow._plan, 1))
    import ctypes as C

    ms = (C.c_double * 16)()
    cnt = C.c_int64()
    lib.sopht_ns3d_stage_times(flow._plan, ms, 16, C.byref(cnt), 1)  # reset
    inter.check_support_in_domain()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(flow, inter, U)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    nst = lib.sopht_ns3d_stage_times(flow._plan, ms, 16, C.byref(cnt), 0)
    names = [lib.sopht_ns3d_stage_name(flow._plan, i).decode() for i in range(nst)]
    avg_ms = {names[i]: ms[i] / max(cnt.value, 1) for i in range(nst)}
    passes, zjobs = L.schedule("sopht_ns3d_poisson_schedule", flow._plan)
    s1_vec, s1_chunk, s1_cols = L.schedule("sopht_ns3d_s1_schedule", flow._plan)
    inter.check_support_in_domain()
    if not torch.isfinite(flow.vorticity_field).all():
        print("non-finite vorticity after the timed steps", file=sys.stderr)
        return 1

    cells = n**3
    ms_per_step = 1e3 * elapsed / args.steps
    value = cells * args.steps / elapsed
    word = np.dtype(real_t).itemsize
    words = stage_words(n, n, n)
    dom = max(avg_ms, key=avg_ms.get)
    alg_bytes = words[dom] * cells * word
    achieved = alg_bytes / (avg_ms[dom] * 1e-3) / 1e9
    traffic = hbm_traffic(dom) if (n == 256 and real_t == np.float32) else None
    roof = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "kernel": dom,
        "kernel_ms": round(avg_ms[dom], 4),
        "algorithmic_bytes": int(alg_bytes),
    }
    step_gbs = STEP_WORDS_PER_CELL * word * cells / (ms_per_step * 1e-3) / 1e9
    kernels_gbs = {
        k: round(words[k] * cells * word / (avg_ms[k] * 1e-3) / 1e9, 1) for k in avg_ms if avg_ms[k] > 0
    }
    batched = None
    if args.batch_steps > 0:
        del flow, inter
        torch.cuda.empty_cache()
        batched = batched_vs_sequential(args.batch_steps, max(2, args.batch_steps // 4))
    adjoint = None
    if args.adjoint_steps > 0:
        torch.cuda.empty_cache()
        adjoint = adjoint_vs_forward(steps=args.adjoint_steps)
    fastdiag = None
    if args.fastdiag_reps > 0:
        torch.cuda.empty_cache()
        fastdiag = fastdiag_solve(reps=args.fastdiag_reps, cpu=args.cpu_steps > 0)
    rod = None
    if args.rod_steps > 0:
        torch.cuda.empty_cache()
        rod = rod_coupled(steps=args.rod_steps)
    sizes = None
    if args.sizes_steps > 0:
        torch.cuda.empty_cache()
        sizes = other_sizes(args.sizes_steps)
    filtered = None
    if args.filter_steps > 0:
        torch.cuda.empty_cache()
        filtered = filtered_step(args.filter_steps)
    factories = None
    if args.factory_reps > 0:
        torch.cuda.empty_cache()
        factories = standalone_factories(reps=args.factory_reps)
    cpu = None
    if args.cpu_steps > 0:
        try:
            cpu = cpu_baseline(n, real_t, args.cpu_steps)
        except Exception as e:  # the baseline is reported, never required
            cpu = {"value": None, "unit": "cell-updates/s", "cores": None, "kind": "port", "sample": f"failed: {e}"}
    out = {
        "metric": "3D IB flow cell-updates/s at 256^3 fp32 (% HBM peak); batched vs plain; adj/fwd",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if real_t == np.float32 else "f64",
        "data": "synthetic (seeded random-phase vorticity, Fibonacci sphere)",
        "config": {
            "workload": WORKLOAD if n == 256 and real_t == np.float32 else f"coupled IB flow step, {n}^3 {args.precision}",
            "grid": [n, n, n],
            "precision": args.precision,
            "lagrangian_nodes": 3000,
            "parallelism": "single GPU",
        },
        "step_hbm_frac_76w": round(step_gbs / HBM_PEAK_GBS, 4),
        "stage_ms": {k: round(v, 4) for k, v in avg_ms.items()},
        "stage_gbs": kernels_gbs,
        "poisson_schedule": {"passes": passes, "zpass_jobs": zjobs},
        "s1_sweep": {"vector": s1_vec, "zchunk": s1_chunk, "cols": s1_cols},
        "roofline": roof,
        "batched": batched,
        "adjoint": adjoint,
        "fast_diagonalisation": fastdiag,
        "rod_coupling": rod,
        "other_sizes": sizes,
        "filtered_step": filtered,
        "standalone_factories": factories,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
