"""Shared pytest setup: repo root on sys.path, the `gpu` marker, helpers."""
import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.build()
    return oracle


def has_gpu() -> bool:
    try:
        import torch  # noqa: F401  (device count without initialising HIP)
        return torch.cuda.device_count() > 0
    except Exception:
        return False


def scaled_err(a, b, scale):
    """max_i |a_i - b_i| / scale_i over vector rows."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.abs(a - b)
    if d.ndim == 2:
        d = d.max(axis=1)
    return float((d / np.maximum(scale, 1e-300)).max())
