"""Build the gfx950 HIP library ``sopht_amd/libsopht_amd.so`` in-tree.

Each ``csrc/*.hip`` is compiled to an object with hipcc in parallel, then linked into one
shared library exposing the C ABI declared in ``include/sopht_amd.h``. The checkpoint file
layer ``csrc/h5io.c`` (host C over the HDF5 C library, ``include/sopht_amd_h5.h``) is built
with gcc into ``libsopht_amd_h5.so`` beside it, so the GPU library does not load libhdf5.
Run as ``python -m sopht_amd.build`` (or through ``__graft_entry__.build()``).
"""

from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libsopht_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops is disabled on purpose: packed FP32 forms are kept out of the code.
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-I",
    str(ROOT / "include"),
    "-Wno-unused-result",
]


# rocBLAS runs the dense eigenbasis GEMMs of the fast-diagonalisation Poisson solver.
ROCM_LIB = "/opt/rocm/lib"
LINK_LIBS = ["-L", ROCM_LIB, f"-Wl,-rpath,{ROCM_LIB}", "-lrocblas"]


# HDF5 1.10 C library and headers of the image (the HDF5 that SophT's h5py wraps).
HDF5_PREFIX = Path(os.environ.get("SOPHT_HDF5_PREFIX", "/opt/conda"))
H5_LIB = PKG / "libsopht_amd_h5.so"
CC = os.environ.get("CC", "gcc")


def build_h5(verbose: bool = True) -> Path:
    """Compile the checkpoint file layer (plain C, no C++ runtime) against libhdf5."""
    tmp = H5_LIB.with_suffix(".so.tmp")
    lib_dir = HDF5_PREFIX / "lib"
    cmd = [CC, "-std=c11", "-O2", "-Wall", "-fPIC", "-shared", "-idirafter", str(HDF5_PREFIX / "include"),
           str(CSRC / "h5io.c"), "-o", str(tmp), "-L", str(lib_dir), f"-Wl,-rpath,{lib_dir}", "-lhdf5"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gcc failed for h5io.c:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, H5_LIB)
    if verbose:
        print(f"built {H5_LIB.relative_to(ROOT)}", file=sys.stderr)
    return H5_LIB


def _compile(src: Path) -> Path:
    obj = OBJ / (src.stem + ".o")
    cmd = [HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    OBJ.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    jobs = min(len(sources), max(1, min(8, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs),
           *LINK_LIBS]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB.relative_to(ROOT)} from {len(sources)} sources", file=sys.stderr)
    build_h5(verbose)
    return LIB


if __name__ == "__main__":
    build()
