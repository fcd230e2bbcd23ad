"""Precision selection, same contract as ``sopht/utils/precision.py:6-19``."""

import numpy as np


def get_real_t(precision: str = "single") -> type:
    """``"single"`` -> ``np.float32``, ``"double"`` -> ``np.float64``; anything else is a
    ``ValueError`` (``sopht/utils/precision.py:6-13``)."""
    if precision == "single":
        return np.float32
    if precision == "double":
        return np.float64
    msg = "Precision argument must be single or double"
    raise ValueError(msg)


def get_test_tol(precision: str = "single") -> float:
    """``1e3 * eps(real_t)`` (``sopht/utils/precision.py:16-19``)."""
    real_t = get_real_t(precision=precision)
    return real_t(1e3) * np.finfo(real_t).eps
