"""Restart from the latest saved snapshot (``sopht/utils/restart_sim.py:11-37``)."""

from __future__ import annotations

import logging
from pathlib import Path

from .io import IO

logger = logging.getLogger(__name__)


def restart_simulation(restart_simulator, io: IO, rod_io: IO | None, forcing_io: IO | None,
                       restart_dir: str, h5_dir: str | None = None) -> float:
    """Load the flow (``sopht_NNNN.h5``), rod (``rod_NNNN.h5``) and forcing-grid
    (``forcing_grid_NNNN.h5``) snapshots with the highest NNNN and return the flow time.

    As in the reference, the snapshots are looked up in the working directory (``h5_dir``
    overrides that) and ``restart_simulator``'s PyElastica state is loaded from
    ``restart_dir`` and must carry the same time. Rigid immersed bodies have no PyElastica
    state: pass ``restart_simulator=None`` and ``rod_io=None`` to restore the flow and the
    virtual-boundary state (the forcing grid's position mismatch) alone.
    """
    base = Path(h5_dir) if h5_dir is not None else Path.cwd()
    iter_num = [int(p.stem.split("_")[-1]) for p in base.glob("sopht_*.h5")]
    if len(iter_num) == 0:
        raise FileNotFoundError("There is no file to load in the directory.")
    latest = max(iter_num)
    curr_time = io.load(h5_file_name=str(base / f"sopht_{latest:04d}.h5"))
    if rod_io is not None:
        rod_io.load(h5_file_name=str(base / f"rod_{latest:04d}.h5"))
    if forcing_io is not None:
        forcing_io.load(h5_file_name=str(base / f"forcing_grid_{latest:04d}.h5"))
    if restart_simulator is not None:
        import elastica as ea  # the rod state lives in PyElastica's own restart files

        rod_time = ea.load_state(restart_simulator, restart_dir, True)
        if curr_time != rod_time:
            raise ValueError("Simulation time of the flow is not matched with the Elastica, check your inputs!")
    logger.info("sopht_%04d.h5 has been loaded", latest)
    return curr_time
