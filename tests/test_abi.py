"""The C ABI library loads and exports every entry point ``include/*.h`` declares; the
Python bindings cover all of them; no compute runs here (no GPU needed)."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "sopht_amd" / "libsopht_amd.so"
H5_LIB = ROOT / "sopht_amd" / "libsopht_amd_h5.so"
# each header declares the entry points of one shared library
LIBRARY_OF = {"sopht_amd.h": LIB, "sopht_amd_h5.h": H5_LIB}


def declared_symbols(header="sopht_amd.h"):
    text = (ROOT / "include" / header).read_text()
    return set(re.findall(r"^(?:const\s+char\s*\*|int64_t|int)\s+(sopht_\w+)\s*\(", text, re.M))


def test_every_header_has_a_library():
    assert {h.name for h in (ROOT / "include").glob("*.h")} == set(LIBRARY_OF)


@pytest.mark.skipif(not H5_LIB.exists(), reason="library not built (python -m sopht_amd.build)")
def test_checkpoint_library_exports_and_bindings():
    from sopht_amd.utils import io

    names = declared_symbols("sopht_amd_h5.h")
    assert {"sopht_h5_file_create", "sopht_h5_dataset_write", "sopht_h5_list"} <= names
    lib = ctypes.CDLL(str(H5_LIB))
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert names == set(io.SIGNATURES)
    io._h5()  # sets argtypes for every symbol; raises if one is missing


def test_header_declares_entry_points():
    names = declared_symbols()
    assert {"sopht_ns3d_step", "sopht_poisson3d_vector_solve", "sopht_lag_to_eul_spread_3d"} <= names


@pytest.mark.skipif(not LIB.exists(), reason="library not built (python -m sopht_amd.build)")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.sopht_version() >= 1


@pytest.mark.skipif(not LIB.exists(), reason="library not built (python -m sopht_amd.build)")
def test_python_bindings_cover_header():
    from sopht_amd import _lib

    assert declared_symbols() == set(_lib.SIGNATURES)
    _lib.lib()  # sets argtypes for every symbol; raises if one is missing


@pytest.mark.skipif(not LIB.exists(), reason="library not built (python -m sopht_amd.build)")
def test_error_paths_need_no_gpu():
    """Argument validation happens on the host before any HIP call."""
    from sopht_amd import _lib

    lib = _lib.lib()
    plan = ctypes.c_void_p()
    rc = lib.sopht_poisson3d_create(0, 10, 10, 10, 1.0, None, ctypes.byref(plan))
    assert rc == _lib.SOPHT_ERR_UNSUPPORTED  # 2n = 20 has a factor 5
    assert b"unsupported Poisson grid size" in lib.sopht_last_error()
    with pytest.raises(ValueError):
        _lib.check(rc)
    assert lib.sopht_poisson3d_create(7, 16, 16, 16, 1.0, None, ctypes.byref(plan)) == _lib.SOPHT_ERR_INVALID
    st = (ctypes.c_int64 * 3)(1, 1, 1)
    assert lib.sopht_penalise_field_boundary_3d(0, None, st, 8, 8, 8, -1, None, None) == _lib.SOPHT_ERR_INVALID
    fd = ctypes.c_void_p()
    assert lib.sopht_fastdiag3d_create(0, 1, 8, 8, 0.1, None, ctypes.byref(fd)) == _lib.SOPHT_ERR_INVALID
    assert lib.sopht_fastdiag3d_create(0, 8, 8, 8, -0.1, None, ctypes.byref(fd)) == _lib.SOPHT_ERR_INVALID
    assert lib.sopht_fastdiag3d_create(3, 8, 8, 8, 0.1, None, ctypes.byref(fd)) == _lib.SOPHT_ERR_INVALID
    ns = ctypes.c_void_p()
    assert lib.sopht_ns3d_create_with_poisson_solver(
        0, 8, 8, 8, 1.0, 1e-2, 1.0, 2, 0, 0, 2, None, ctypes.byref(ns)) == _lib.SOPHT_ERR_INVALID


def test_product_does_not_import_oracle():
    for p in (ROOT / "sopht_amd").rglob("*.py"):
        src = p.read_text()
        assert "import oracle" not in src and "from oracle" not in src, p


def test_reference_validation_errors():
    """Factory-time ValueErrors of the reference (no device work)."""
    import numpy as np

    from sopht_amd.numeric import eulerian_grid_ops as spne
    from sopht_amd.utils import get_real_t

    with pytest.raises(ValueError):
        get_real_t("half")
    with pytest.raises(ValueError):
        spne.gen_set_fixed_val_pyst_kernel_3d(np.float32, field_type="tensor")
    with pytest.raises(ValueError):
        spne.gen_penalise_field_boundary_pyst_kernel_3d(-1, 0.1, None, None, None, np.float32)
    with pytest.raises(ValueError):
        spne.gen_set_fixed_val_at_boundaries_pyst_kernel_3d(np.float32, width=0)


def test_laplacian_filter_factory_validation():
    """gen_laplacian_filter_kernel_3d's factory-time ValueErrors (laplacian_filter_3d.py:40-50,
    :168-174) need no device."""
    import numpy as np

    from sopht_amd.numeric.eulerian_grid_ops import gen_laplacian_filter_kernel_3d

    for kw in ({"filter_order": -1}, {"filter_order": 1.5}, {"filter_type": "gaussian"},
               {"filter_flux_buffer_boundary_width": 0}):
        args = dict(filter_order=1, filter_flux_buffer=None, field_buffer=None, real_t=np.float32)
        args.update(kw)
        with pytest.raises(ValueError):
            gen_laplacian_filter_kernel_3d(**args)


def test_fast_diag_bc_type_validation():
    """FastDiagPoissonSolver3D defines homogenous_neumann_along_xyz only (FastDiagPoissonSolver3D.py:22)."""
    import numpy as np

    from sopht_amd.numeric.eulerian_grid_ops import FastDiagPoissonSolver3D

    with pytest.raises(ValueError):
        FastDiagPoissonSolver3D(8, 8, 8, dx=0.1, real_t=np.float32, bc_type="dirichlet")
