"""CPU oracle for the 3D unbounded vorticity flow step with virtual-boundary IB coupling.

TEST INFRASTRUCTURE ONLY. This package is a plain numpy/scipy restatement of the
reference algorithm (SophT, ``sopht/numeric`` and ``sopht/simulator``), written from
the reference's source files read as text. It is the checker for the HIP product path:

* only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg
  may import it;
* nothing in ``sopht_amd`` imports it, and the product path never routes through it.

Every function cites the reference file:line it restates. How each part is pinned
(closed forms, brute force, invariants) is recorded in the module headers; the tests
that do the pinning live in ``tests/test_oracle_*.py``.
"""

from . import flow, ib, poisson, stencils  # noqa: F401
