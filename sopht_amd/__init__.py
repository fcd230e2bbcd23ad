"""sopht_amd: MI355X (gfx950) backend for SophT's 3D unbounded vorticity flow step with
virtual-boundary immersed-boundary coupling.

Subpackages mirror the reference layout for this path: ``numeric`` (grid stencils, the
free-space Poisson solver, Eulerian-Lagrangian transfer and virtual-boundary forcing),
``simulator`` (the 3D flow simulator and immersed-body glue) and ``utils``. Compute runs in
the hand-written HIP library ``libsopht_amd.so`` (C ABI: ``include/sopht_amd.h``) built by
``python -m sopht_amd.build``; importing a kernel factory without it raises ``ImportError``.
"""

__version__ = "0.1.0"
