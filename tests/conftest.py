"""Common fixtures (the reference's ``tests/conftest.py:10-24``) and the ``gpu`` marker."""

import os
import sys
from pathlib import Path

import pytest
from numpy.random import Generator, default_rng

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built library")


@pytest.fixture
def rng() -> Generator:
    """Random generator with a fixed seed, as the reference's fixture."""
    return default_rng(seed=42)


@pytest.fixture
def max_cpu_count() -> int:
    if hasattr(os, "sched_getaffinity"):
        return len(os.sched_getaffinity(0))
    return os.cpu_count() or 1


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture
def dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda", 0)
