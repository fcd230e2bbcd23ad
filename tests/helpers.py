"""Shared helpers for the GPU parity tests."""

import numpy as np

PREC = {"single": np.float32, "double": np.float64}


def tol(real_t):
    """The reference's test tolerance ``1e3 * eps`` (``sopht/utils/precision.py:16-19``)."""
    return real_t(1e3) * np.finfo(real_t).eps


def to_dev(a, dev):
    import torch

    return torch.as_tensor(np.ascontiguousarray(a)).to(dev)


def to_host(t):
    return t.detach().cpu().numpy()


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0)


def rel_linf(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).max()
    return np.abs(a - b).max() / (den if den > 0 else 1.0)
