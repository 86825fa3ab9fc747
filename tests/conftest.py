"""Shared fixtures.  GPU tests are marked `gpu` and need an MI355X (gfx950)."""
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD MI355X GPU (gfx950) and the built HIP library")


@pytest.fixture(scope="session")
def be():
    from fluxion_amd.hip import HipBackend
    return HipBackend(0)


def gaussian_ic(n, h):
    """Gaussian centred at (0.5, 0.5), sigma 0.1 (crates/runtime/tests/solvers.rs:23-38)."""
    x = np.arange(n) * h
    X, Y = np.meshgrid(x, x)
    r2 = (X - 0.5) ** 2 + (Y - 0.5) ** 2
    return np.exp(-r2 * (1.0 / (2.0 * 0.1 * 0.1))).ravel()


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
