"""pytest configuration: the `gpu` marker and import paths.

GPU tests (marked `gpu`) call the HIP engine through its C ABI and use the CPU
oracle (oracle/) as the checker. Everything else runs on the CPU.
"""
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT / "n-body_amd", ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def ora():
    import oracle

    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def nb():
    import nbody_amd

    nbody_amd.load()
    return nbody_amd


def rel_err(a, b):
    """Per-body relative error |a - b| / |b| of (n, >=3) arrays."""
    a = np.asarray(a, np.float64)[:, :3]
    b = np.asarray(b, np.float64)[:, :3]
    den = np.linalg.norm(b, axis=1)
    den = np.where(den > 0, den, 1.0)
    return np.linalg.norm(a - b, axis=1) / den
