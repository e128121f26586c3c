"""Shared test setup.

Markers: ``gpu`` tests need an MI355X (run with ``-m gpu``); everything else
runs on the CPU.  The CPU oracle (oracle/) is the checker; the HIP library
(giant-impact_amd/lib/libgiant_impact.so) is what the GPU tests exercise,
through its C ABI.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent
for p in (REPO, REPO / "giant-impact_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.load()
    return O


@pytest.fixture(scope="session")
def gi():
    import gi_amd

    gi_amd.load()
    return gi_amd


@pytest.fixture(scope="session")
def gpu(gi):
    """The HIP library with a visible device (GPU tests only)."""
    if gi.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests must run on the MI355X box")
    return gi


def f4(n: int) -> np.ndarray:
    return np.zeros((n, 4), dtype=np.float32)
