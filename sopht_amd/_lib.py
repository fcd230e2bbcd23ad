"""ctypes binding of ``libsopht_amd.so`` (the C ABI in ``include/sopht_amd.h``).

The library is loaded from this package directory (or from the path in the environment
variable ``SOPHT_AMD_LIB``); there is no fallback. A missing library raises ``ImportError``
at first use, naming the build command. Device buffers are
passed as raw pointers (``tensor.data_ptr()``) of torch tensors on the current HIP device;
every launch goes on torch's current stream.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(os.environ.get("SOPHT_AMD_LIB") or Path(__file__).resolve().parent / "libsopht_amd.so")

SOPHT_F32 = 0
SOPHT_F64 = 1
SOPHT_ERR_INVALID = -1
SOPHT_ERR_UNSUPPORTED = -2
SOPHT_ERR_HIP = -3
SOPHT_ERR_OUT_OF_DOMAIN = -4

_vp = C.c_void_p
_i64 = C.c_int64
_i = C.c_int
_d = C.c_double
_p3 = _vp * 3
_pi64 = C.POINTER(C.c_int64)
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)



class BatchMember(C.Structure):
    """``sopht_batch_member`` (include/sopht_amd.h)."""

    _fields_ = [
        ("nz", C.c_int64), ("ny", C.c_int64), ("nx", C.c_int64),
        ("x_range", C.c_double), ("nu", C.c_double), ("rho", C.c_double), ("cfl", C.c_double),
        ("width", C.c_int), ("with_forcing", C.c_int), ("with_free_stream", C.c_int), ("kernel_type", C.c_int),
        ("free_stream", C.c_double * 3),
        ("time", C.c_double),
        ("vorticity", C.c_void_p), ("velocity", C.c_void_p), ("psi", C.c_void_p), ("forcing", C.c_void_p),
        ("n_lag", C.c_int64),
        ("dx", C.c_double), ("shift", C.c_double), ("k", C.c_double), ("c", C.c_double),
        ("X", C.c_void_p), ("V", C.c_void_p),
        ("idx", C.c_void_p), ("support", C.c_void_p), ("weights", C.c_void_p), ("lag_u", C.c_void_p),
        ("dV", C.c_void_p), ("dX", C.c_void_p), ("Fl", C.c_void_p),
    ]


# name -> argtypes (restype int unless listed in _RESTYPES)
SIGNATURES = {
    "sopht_last_error": [],
    "sopht_version": [],
    "sopht_curl_3d": [_i, _p3, _pi64, _p3, _pi64, _i64, _i64, _i64, _d, _i, _vp],
    "sopht_update_vorticity_from_velocity_forcing_3d": [_i, _p3, _pi64, _p3, _pi64, _i64, _i64, _i64, _d, _vp],
    "sopht_update_vorticity_from_penalised_velocity_3d": [
        _i, _p3, _pi64, _p3, _pi64, _p3, _pi64, _i64, _i64, _i64, _d, _vp],
    "sopht_elementwise_cross_product_3d": [_i, _p3, _pi64, _p3, _pi64, _p3, _pi64, _i64, _i64, _i64, _vp],
    "sopht_diffusion_flux_3d": [_i, _vp, _pi64, _vp, _pi64, _i64, _i64, _i64, _d, _i, _vp],
    "sopht_diffusion_timestep_euler_forward_3d": [_i, _vp, _pi64, _vp, _pi64, _i64, _i64, _i64, _d, _vp],
    "sopht_penalise_field_boundary_3d": [_i, _vp, _pi64, _i64, _i64, _i64, _i, _pd, _vp],
    "sopht_divergence_3d": [_i, _vp, _pi64, _p3, _pi64, _i64, _i64, _i64, _d, _i, _vp],
    "sopht_set_fixed_val_3d": [_i, _vp, _pi64, _i64, _i64, _i64, _d, _vp],
    "sopht_add_fixed_val_3d": [_i, _vp, _pi64, _vp, _pi64, _i64, _i64, _i64, _d, _vp],
    "sopht_elementwise_saxpby_3d": [_i, _vp, _pi64, _vp, _pi64, _vp, _pi64, _i64, _i64, _i64, _d, _d, _vp],
    "sopht_elementwise_sum_3d": [_i, _vp, _pi64, _vp, _pi64, _vp, _pi64, _i64, _i64, _i64, _vp],
    "sopht_elementwise_copy_3d": [_i, _vp, _pi64, _vp, _pi64, _i64, _i64, _i64, _vp],
    "sopht_set_fixed_val_at_boundaries_3d": [_i, _vp, _pi64, _i64, _i64, _i64, _i, _d, _vp],
    "sopht_elementwise_complex_product_3d": [_i, _vp, _pi64, _vp, _pi64, _vp, _pi64, _i64, _i64, _i64, _vp],
    "sopht_advection_flux_conservative_eno3_3d": [
        _i, _vp, _pi64, _vp, _pi64, _p3, _pi64, _i64, _i64, _i64, _d, _vp],
    "sopht_advection_flux_conservative_eno3_adjoint_3d": [
        _i, _vp, _pi64, _vp, _pi64, _vp, _pi64, _vp, _pi64, _p3, _pi64, _i64, _i64, _i64, _d, _vp],
    "sopht_passive_transport_step_3d": [_i, _i, _p3, _p3, _p3, _i64, _i64, _i64, _d, _d, _vp],
    "sopht_velocity_abs_sum_max_3d": [_i, _p3, _i64, _vp, _vp, _i64, _vp, _vp],
    "sopht_vorticity_stretching_flux_3d": [_i, _p3, _pi64, _p3, _pi64, _p3, _pi64, _i64, _i64, _i64, _d, _vp],
    "sopht_poisson3d_create": [_i, _i64, _i64, _i64, _d, _vp, C.POINTER(_vp)],
    "sopht_poisson3d_solve": [_vp, _vp, _pi64, _vp, _pi64, _vp],
    "sopht_poisson3d_vector_solve": [_vp, _p3, _pi64, _p3, _pi64, _vp],
    "sopht_poisson3d_destroy": [_vp],
    "sopht_poisson3d_scratch_bytes": [_vp],
    "sopht_poisson3d_schedule": [_vp, _pi, _pi],
    "sopht_fastdiag3d_create": [_i, _i64, _i64, _i64, _d, _vp, C.POINTER(_vp)],
    "sopht_fastdiag3d_solve": [_vp, _vp, _pi64, _vp, _pi64, _vp],
    "sopht_fastdiag3d_vector_solve": [_vp, _p3, _pi64, _p3, _pi64, _vp],
    "sopht_fastdiag3d_eigenbasis": [_vp, _i, _pd, _pd],
    "sopht_fastdiag3d_destroy": [_vp],
    "sopht_lag_support_3d": [_i, _vp, _vp, _vp, _i64, _d, _d, _vp],
    "sopht_lag_interp_weights_3d": [_i, _i, _vp, _vp, _i64, _d, _vp],
    "sopht_eul_to_lag_interp_3d": [_i, _i, _vp, _p3, _pi64, _i64, _i64, _i64, _vp, _vp, _i64, _d, _vp],
    "sopht_spread3d_create": [_i64, C.POINTER(_vp)],
    "sopht_spread3d_destroy": [_vp],
    "sopht_lag_to_eul_spread_3d": [_i, _i, _vp, _p3, _pi64, _i64, _i64, _i64, _vp, _vp, _vp, _i64, _vp],
    "sopht_spread3d_check": [_vp, _vp],
    "sopht_spread3d_reuse_plan": [_vp],
    "sopht_lag_to_eul_spread_3d_noalloc": [_i, _i, _p3, _pi64, _i64, _i64, _i64, _vp, _vp, _vp, _i64, _vp, _vp],
    "sopht_vbf_lag_force_3d": [
        _i, _i, _p3, _pi64, _i64, _i64, _i64, _vp, _vp, _i64, _d, _d, _d, _d,
        _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "sopht_vbf_time_step_3d": [_i, _vp, _vp, _i64, _d, _vp],
    "sopht_ns3d_create": [_i, _i64, _i64, _i64, _d, _d, _d, _i, _i, _i, _vp, C.POINTER(_vp)],
    "sopht_ns3d_create_with_poisson_solver": [
        _i, _i64, _i64, _i64, _d, _d, _d, _i, _i, _i, _i, _vp, C.POINTER(_vp)],
    "sopht_ns3d_destroy": [_vp],
    "sopht_ns3d_step": [_vp, _vp, _vp, _vp, _vp, _vp, _d, _pd, _vp],
    "sopht_ns3d_recover_velocity": [_vp, _vp, _vp, _vp, _pd, _vp],
    "sopht_ns3d_max_abs_velocity_sum": [_vp, _vp, _vp, _vp],
    "sopht_ns3d_last_velocity_max": [_vp, _vp, _vp, _vp],
    "sopht_ns3d_set_filter": [_vp, _i, _i],
    "sopht_laplacian_filter_3d": [_i, _i, _i, _vp, _pi64, _i64, _i64, _i64, _vp, _pi64, _vp, _pi64, _i, _vp],
    "sopht_ns3d_enable_timing": [_vp, _i],
    "sopht_ns3d_stage_times": [_vp, _pd, _i, _pi64, _i],
    "sopht_ns3d_stage_name": [_vp, _i],
    "sopht_ns3d_poisson_schedule": [_vp, _pi, _pi],
    "sopht_ns3d_s1_schedule": [_vp, _pi, _pi, _pi],
    "sopht_ns3d_batch_class_schedule": [_vp, _i, _pi64, _pi, _pi, _pi],
    "sopht_ns3d_batch_create": [_i, _i, C.POINTER(BatchMember), _vp, C.POINTER(_vp)],
    "sopht_ns3d_batch_step": [_vp, _i, _i, _vp],
    "sopht_ns3d_batch_state": [_vp, _pd, _pd, _vp],
    "sopht_ns3d_batch_destroy": [_vp],
    "sopht_ns3d_trajectory_step": [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _pd, _vp],
    "sopht_ns3d_adjoint_step": [_vp, _vp, _vp, _d, _vp, _vp, _vp, _vp],
    "sopht_vbf_adjoint_3d": [
        _i, _vp, _d, _vp, _i64, _i64, _i64, _vp, _vp, _i64, _d, _d, _d, _d, _d, _vp,
        _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
}
_RESTYPES = {
    "sopht_last_error": C.c_char_p,
    "sopht_poisson3d_scratch_bytes": _i64,
    "sopht_ns3d_stage_name": C.c_char_p,
}

_lib = None


def lib():
    """The loaded library (loaded once). Raises ``ImportError`` if it was never built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            msg = (
                f"{LIB_PATH} is missing: build the HIP library first "
                "(python -m sopht_amd.build). There is no CPU fallback."
            )
            raise ImportError(msg)
        handle = C.CDLL(os.fspath(LIB_PATH))
        for name, argtypes in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = argtypes
            fn.restype = _RESTYPES.get(name, C.c_int)
        _lib = handle
    return _lib


def check(rc: int) -> None:
    """Map a SOPHT_ERR_* return code to the reference's exception types."""
    if rc == 0:
        return
    msg = lib().sopht_last_error().decode(errors="replace")
    if rc in (SOPHT_ERR_INVALID, SOPHT_ERR_UNSUPPORTED):
        raise ValueError(msg)
    if rc == SOPHT_ERR_OUT_OF_DOMAIN:
        raise IndexError(msg)
    raise RuntimeError(msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


# ----------------------------------------------------------------------------- helpers


def dtype_tag(real_t) -> int:
    real_t = np.dtype(real_t).type
    if real_t is np.float32:
        return SOPHT_F32
    if real_t is np.float64:
        return SOPHT_F64
    msg = "Invalid real type"
    raise ValueError(msg)


def torch_dtype(real_t):
    import torch

    return torch.float32 if dtype_tag(real_t) == SOPHT_F32 else torch.float64


def stream_ptr():
    """The current HIP stream of the current device (torch's raw stream handle: building a
    ``torch.cuda.Stream`` object per call costs a few microseconds of host time per launch)."""
    import torch

    return _vp(torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice()))


def strides3(t) -> C.Array:
    """Element strides (z, y, x) of a scalar field tensor (or of one vector component)."""
    s = t.stride()[-3:]
    return (C.c_int64 * 3)(*s)


def ptr(t):
    return _vp(t.data_ptr()) if t is not None else _vp(0)


def comps(v):
    """Three component pointers of a (3, nz, ny, nx) tensor (from its base pointer and first
    stride, without building the component views)."""
    base, step = v.data_ptr(), v.stride(0) * v.element_size()
    return _p3(base, base + step, base + 2 * step)


def require_device_tensor(t, real_t, name, ndim=None):
    import torch

    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        msg = f"{name} must be a torch tensor on the HIP device"
        raise TypeError(msg)
    if t.dtype != torch_dtype(real_t):
        msg = f"{name} has dtype {t.dtype}, expected {torch_dtype(real_t)}"
        raise TypeError(msg)
    if ndim is not None and t.dim() != ndim:
        msg = f"{name} must have {ndim} dimensions, got shape {tuple(t.shape)}"
        raise ValueError(msg)


def written(*tensors) -> None:
    """Bump torch's version counter of tensors a library call wrote through raw pointers.

    Writes through ``data_ptr()`` are invisible to torch; callers that cache results keyed
    on a tensor's version (the simulator's fused velocity maxima) rely on this to see them.
    Views share their base's counter, so marking a component marks the whole field."""
    import torch

    live = [t for t in tensors if t is not None]
    if live:
        torch.autograd.graph.increment_version(live)


def schedule(fn: str, handle) -> tuple[int, ...]:
    """The integers a plan reports: (passes, zpass_jobs) of a Poisson plan
    (``sopht_poisson3d_schedule``) or of a step plan (``sopht_ns3d_poisson_schedule``), or
    (vector, zchunk, cols) of a step plan's first sweep (``sopht_ns3d_s1_schedule``)."""
    outs = [C.c_int(0) for _ in range(len(SIGNATURES[fn]) - 1)]
    check(getattr(lib(), fn)(handle, *(C.byref(o) for o in outs)))
    return tuple(o.value for o in outs)
