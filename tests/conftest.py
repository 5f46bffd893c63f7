import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def small_panel():
    from hdfe_stream_amd.simulate import akm_panel
    return akm_panel(2000, 200, seed=1)


@pytest.fixture(scope="session")
def c1_panel():
    """The 20,000-worker x 8-year, 2,000-firm AKM panel (seed 1)."""
    from hdfe_stream_amd.simulate import akm_panel
    return akm_panel(20_000, 2_000, seed=1)
