"""GPU parity of the free-space Poisson solve (pruned FFT passes) against the oracle and
against the brute-force convolution; sizes include the reference test (n = 16, x_range 2),
non-cubic grids, radix-3 lengths and strided operands. FP64 agreement is required to
relative L2/Linf <= 1e-10; FP32 to the reference's 1e3*eps scaled by the solution size."""

import numpy as np
import pytest

from helpers import PREC, rel_l2, rel_linf, to_dev, to_host
from oracle import poisson as P

pytestmark = pytest.mark.gpu


def solver_cls():
    from sopht_amd.numeric.eulerian_grid_ops import UnboundedPoissonSolverPYFFTW3D

    return UnboundedPoissonSolverPYFFTW3D


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("shape,x_range", [((16, 16, 16), 2.0), ((8, 12, 16), 1.0), ((24, 24, 24), 1.0), ((6, 4, 8), 0.7),
                                           ((256, 8, 8), 1.0), ((256, 6, 16), 1.5)])
def test_vector_solve_matches_oracle(dev, precision, shape, x_range, rng):
    real_t = PREC[precision]
    rhs = rng.standard_normal((3, *shape)).astype(real_t)
    ref = np.zeros_like(rhs, dtype=np.float64)
    P.UnboundedPoissonSolver3D(*shape, x_range=x_range, real_t=np.float64).vector_field_solve(ref, rhs.astype(np.float64))
    s = solver_cls()(*shape, x_range=x_range, real_t=real_t)
    out = to_dev(np.zeros_like(rhs), dev)
    s.vector_field_solve(solution_vector_field=out, rhs_vector_field=to_dev(rhs, dev))
    got = to_host(out)
    bound = 1e-10 if real_t == np.float64 else 2e-5
    assert rel_l2(got, ref) <= bound, rel_l2(got, ref)
    assert rel_linf(got, ref) <= 5 * bound, rel_linf(got, ref)


@pytest.mark.parametrize("precision", ["single", "double"])
def test_scalar_solve_reference_case(dev, precision, rng):
    """test_unbounded_poisson_solver_3d.py:101-125 (n = 16, x_range = 2)."""
    real_t = PREC[precision]
    rhs = rng.standard_normal((16, 16, 16)).astype(real_t)
    ref = np.zeros_like(rhs)
    P.UnboundedPoissonSolver3D(16, 16, 16, x_range=real_t(2.0), real_t=real_t).solve(ref, rhs)
    s = solver_cls()(16, 16, 16, x_range=real_t(2.0), real_t=real_t)
    out = to_dev(np.zeros_like(rhs), dev)
    s.solve(solution_field=out, rhs_field=to_dev(rhs, dev))
    np.testing.assert_allclose(to_host(out), ref, atol=real_t(1e3) * np.finfo(real_t).eps)


def test_brute_force_fp64(dev, rng):
    rhs = rng.standard_normal((6, 6, 6))
    s = solver_cls()(6, 6, 6, x_range=1.3, real_t=np.float64)
    out = to_dev(np.zeros_like(rhs), dev)
    s.solve(out, to_dev(rhs, dev))
    ref = P.brute_force_free_space_solve(rhs, s.dx)
    assert rel_linf(to_host(out), ref) < 1e-12


def test_rhs_untouched_and_strided(dev, rng):
    rhs = rng.standard_normal((3, 16, 16, 16))
    d_rhs = to_dev(rhs, dev)
    big = to_dev(np.zeros((3, 18, 16, 20)), dev)
    view = big[:, 1:17, :, 2:18]
    s = solver_cls()(16, 16, 16, real_t=np.float64)
    s.vector_field_solve(view, d_rhs)
    np.testing.assert_array_equal(to_host(d_rhs), rhs)
    ref = np.zeros_like(rhs)
    P.UnboundedPoissonSolver3D(16, 16, 16).vector_field_solve(ref, rhs)
    assert rel_l2(to_host(view), ref) < 1e-12
    b = to_host(big)
    assert np.all(b[:, 0] == 0) and np.all(b[:, :, :, :2] == 0)


@pytest.mark.parametrize("n", [64, 96])
def test_fp64_larger_grids(dev, n, rng):
    rhs = rng.standard_normal((3, n, n, n))
    s = solver_cls()(n, n, n, real_t=np.float64)
    out = to_dev(np.zeros_like(rhs), dev)
    s.vector_field_solve(out, to_dev(rhs, dev))
    ref = np.zeros_like(rhs)
    P.UnboundedPoissonSolver3D(n, n, n).vector_field_solve(ref, rhs)
    assert rel_l2(to_host(out), ref) < 1e-12
    assert rel_linf(to_host(out), ref) < 1e-11


def test_linearity_at_256(dev):
    """Size-independent property at the benchmark size: solve(a x + b y) = a solve(x) + b solve(y)."""
    import torch

    n = 256
    g = torch.Generator(device=dev).manual_seed(5)
    x = torch.randn((3, n, n, n), device=dev, dtype=torch.float32, generator=g)
    y = torch.randn((3, n, n, n), device=dev, dtype=torch.float32, generator=g)
    s = solver_cls()(n, n, n, real_t=np.float32)
    px, py, pz = (torch.empty_like(x) for _ in range(3))
    s.vector_field_solve(px, x)
    s.vector_field_solve(py, y)
    s.vector_field_solve(pz, 2.0 * x - 0.5 * y)
    err = torch.linalg.vector_norm(pz - (2.0 * px - 0.5 * py)) / torch.linalg.vector_norm(pz)
    assert err.item() < 1e-5


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("shape", [(192, 6, 8), (8, 192, 6), (6, 8, 192), (4, 6, 384), (384, 4, 6)])
def test_solve_on_384_and_768_point_lines(dev, precision, shape, rng):
    """The twelve-point-per-thread plans of the 384- and 768-point lines (the 192^3 workload's
    doubled axes) in every pass, against the oracle."""
    real_t = PREC[precision]
    rhs = rng.standard_normal((3, *shape)).astype(real_t)
    ref = np.zeros_like(rhs, dtype=np.float64)
    P.UnboundedPoissonSolver3D(*shape, x_range=1.0, real_t=np.float64).vector_field_solve(ref, rhs.astype(np.float64))
    s = solver_cls()(*shape, x_range=1.0, real_t=real_t)
    out = to_dev(np.zeros_like(rhs), dev)
    s.vector_field_solve(solution_vector_field=out, rhs_vector_field=to_dev(rhs, dev))
    bound = 1e-10 if real_t == np.float64 else 2e-5
    assert rel_l2(to_host(out), ref) <= bound, rel_l2(to_host(out), ref)


@pytest.mark.parametrize("shape", [(64, 64, 64), (48, 64, 64), (32, 128, 128), (12, 192, 192), (48, 48, 48),
                                   (24, 96, 96)])
def test_three_pass_schedule_matches_oracle(dev, shape, rng):
    """FP32 grids with nx = ny and 2 nx in {96, 128, 192, 256, 384, 512} run the three-pass
    schedule (the x transform fused with the inner sums of the y transform, and the outer
    4-point sums in the z pass); against the FP64 oracle at the five-pass schedule's FP32 bound."""
    rhs = rng.standard_normal((3, *shape)).astype(np.float32)
    ref = np.zeros_like(rhs, dtype=np.float64)
    P.UnboundedPoissonSolver3D(*shape, x_range=1.0, real_t=np.float64).vector_field_solve(ref, rhs.astype(np.float64))
    s = solver_cls()(*shape, x_range=1.0, real_t=np.float32)
    sched = s.fft_schedule
    assert sched["passes"] == 3 and sched["zpass_jobs"] in (1, 2, 3, 6), sched
    out = to_dev(np.zeros_like(rhs), dev)
    s.vector_field_solve(solution_vector_field=out, rhs_vector_field=to_dev(rhs, dev))
    got = to_host(out)
    assert rel_l2(got, ref) <= 2e-5, rel_l2(got, ref)
    assert rel_linf(got, ref) <= 1e-4, rel_linf(got, ref)
    # the scalar solve takes the same passes with one component
    one = to_dev(np.zeros_like(rhs[0]), dev)
    s.solve(solution_field=one, rhs_field=to_dev(rhs[1], dev))
    assert rel_l2(to_host(one), ref[1]) <= 2e-5


@pytest.mark.parametrize("n", [192, 256])
def test_three_pass_against_five_pass(dev, n, monkeypatch):
    """At the workloads' sizes, the three-pass solution against the five-pass one on the same
    FP32 right-hand side (SOPHT_POISSON_PASSES=5 at plan creation selects the five passes)."""
    import torch

    gen = torch.Generator(device=dev).manual_seed(7)
    rhs = torch.randn((3, n, n, n), dtype=torch.float32, device=dev, generator=gen)
    outs = []
    for passes in ("5", None):
        if passes:
            monkeypatch.setenv("SOPHT_POISSON_PASSES", passes)
        else:
            monkeypatch.delenv("SOPHT_POISSON_PASSES", raising=False)
        s = solver_cls()(n, n, n, x_range=1.0, real_t=np.float32)
        out = torch.zeros_like(rhs)
        s.vector_field_solve(solution_vector_field=out, rhs_vector_field=rhs)
        outs.append(out)
        del s
    err = (torch.linalg.vector_norm(outs[1] - outs[0]) / torch.linalg.vector_norm(outs[0])).item()
    print(f"{n}^3 FP32 three-pass vs five-pass rel-L2 {err:.3e}")
    assert err <= 1e-5, err


def test_schedule_choice_and_zpass_jobs(dev, monkeypatch):
    """FP64 plans and grids with nx != ny take five passes; the z pass's jobs per workgroup
    (timed at plan creation, or fixed by SOPHT_POISSON_ZJOBS) do not change the result."""
    import torch

    assert solver_cls()(16, 16, 16, x_range=1.0, real_t=np.float64).fft_schedule == {"passes": 5, "zpass_jobs": 0}
    assert solver_cls()(64, 32, 64, x_range=1.0, real_t=np.float32).fft_schedule["passes"] == 5
    n = 96
    gen = torch.Generator(device=dev).manual_seed(3)
    rhs = torch.randn((3, n, n, n), dtype=torch.float32, device=dev, generator=gen)
    outs = []
    for jobs in ("1", "6", "2"):
        monkeypatch.setenv("SOPHT_POISSON_ZJOBS", jobs)
        s = solver_cls()(n, n, n, x_range=1.0, real_t=np.float32)
        assert s.fft_schedule == {"passes": 3, "zpass_jobs": int(jobs)}
        out = torch.zeros_like(rhs)
        s.vector_field_solve(solution_vector_field=out, rhs_vector_field=rhs)
        outs.append(out)
        del s
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])
