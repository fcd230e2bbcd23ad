"""The allocation-free spreading entry point (``sopht_lag_to_eul_spread_3d_noalloc``) against
the workspace form (``sopht_lag_to_eul_spread_3d``), through the C ABI: bit-identical
Eulerian forcing on two calls in a row, the second with fewer nodes on another grid (the
scratch still holds the first call's plan), FP32 and FP64; the node limit is rejected with
a clear error; an out-of-domain support is reported through the output flag."""

import ctypes as C

import numpy as np
import pytest

from helpers import PREC, to_dev, to_host
from oracle import ib

pytestmark = pytest.mark.gpu


def L():
    from sopht_amd import _lib

    return _lib


def spread_both(dev, real_t, n, X, rng, in_domain=True):
    """Spread random Lagrangian forces at nodes X onto an (n, n, n) grid twice from the same
    start: through a fresh workspace, and through the allocation-free entry point; with
    every support inside the grid, also the oracle's serial loop."""
    import torch

    lib = L()
    m = X.shape[1]
    dx = real_t(1.0 / n)
    idx, d = ib.local_eulerian_grid_support(X, dx, real_t(dx / 2), real_t)
    w, _ = ib.cosine_weights(d, dx, real_t)
    Fl = rng.standard_normal((3, m)).astype(real_t)
    F0 = rng.standard_normal((3, n, n, n)).astype(real_t)
    tag = lib.dtype_tag(real_t)
    lag, wd, idd = to_dev(Fl, dev), to_dev(w, dev), to_dev(idx, dev)
    FA, FB = to_dev(F0, dev), to_dev(F0, dev)
    ws = C.c_void_p()
    lib.call("sopht_spread3d_create", m, C.byref(ws))
    try:
        lib.call("sopht_lag_to_eul_spread_3d", tag, 3, ws, lib.comps(FA), lib.strides3(FA[0]), n, n, n,
                 lib.ptr(lag), lib.ptr(wd), lib.ptr(idd), m, lib.stream_ptr())
        if in_domain:
            lib.call("sopht_spread3d_check", ws, lib.stream_ptr())
    finally:
        lib.lib().sopht_spread3d_destroy(ws)
    flag = torch.full((1,), 7, dtype=torch.int32, device=dev)
    lib.call("sopht_lag_to_eul_spread_3d_noalloc", tag, 3, lib.comps(FB), lib.strides3(FB[0]), n, n, n,
             lib.ptr(lag), lib.ptr(wd), lib.ptr(idd), m, lib.ptr(flag), lib.stream_ptr())
    torch.cuda.synchronize()
    ref = None
    if in_domain:
        ref = F0.copy()
        ib.spread(ref, Fl, w, idx)
    return to_host(FA), to_host(FB), int(flag.item()), ref


@pytest.mark.parametrize("precision", ["single", "double"])
def test_noalloc_spread_matches_workspace_form(dev, precision):
    real_t = PREC[precision]
    rng = np.random.default_rng(3)
    # first call: 3000 nodes on a 48^3 grid; second: 700 clustered nodes on a 32^3 grid
    X1 = 0.2 + 0.6 * rng.random((3, 3000))
    a, b, flag, ref = spread_both(dev, real_t, 48, X1, rng)
    assert flag == 0
    np.testing.assert_array_equal(b, a)
    np.testing.assert_array_equal(b, ref)
    X2 = 0.5 + 0.2 * rng.random((3, 700)) / 32
    a, b, flag, ref = spread_both(dev, real_t, 32, X2, rng)
    assert flag == 0
    np.testing.assert_array_equal(b, a)
    np.testing.assert_array_equal(b, ref)


def test_noalloc_spread_limits_and_out_of_domain(dev):
    import torch

    lib = L()
    real_t = np.float32
    n = 16
    F = torch.zeros((3, n, n, n), dtype=torch.float32, device=dev)
    big = 8193
    lag = torch.zeros((3, big), dtype=torch.float32, device=dev)
    w = torch.zeros((4, 4, 4, big), dtype=torch.float32, device=dev)
    idx = torch.full((3, big), n // 2, dtype=torch.int64, device=dev)
    rc = lib.lib().sopht_lag_to_eul_spread_3d_noalloc(
        lib.dtype_tag(real_t), 3, lib.comps(F), lib.strides3(F[0]), n, n, n, lib.ptr(lag), lib.ptr(w),
        lib.ptr(idx), big, None, lib.stream_ptr())
    assert rc == lib.SOPHT_ERR_INVALID
    assert "8192" in lib.lib().sopht_last_error().decode()
    # a node next to the grid's edge: its support leaves the grid, the flag says so, and the
    # in-domain node is still spread
    rng = np.random.default_rng(5)
    X = np.array([[0.5, 0.01], [0.5, 0.5], [0.5, 0.5]])
    a, b, flag, _ = spread_both(dev, real_t, n, X, rng, in_domain=False)
    assert flag == 1
    np.testing.assert_array_equal(b, a)
    # a following in-domain call starts from a clean flag
    X = np.array([[0.5, 0.45], [0.5, 0.5], [0.5, 0.55]])
    a, b, flag, ref = spread_both(dev, real_t, n, X, rng)
    assert flag == 0
    np.testing.assert_array_equal(b, ref)
