"""Test configuration: the ``gpu`` marker, paths, and shared fixtures."""

from __future__ import annotations

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built liblucem_hip.so")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.build()
    return O


@pytest.fixture(scope="session")
def gpu():
    """The default client on device 0 (the HIP library loaded, no fallback)."""
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu-marked test run without a HIP device")
    from lucem_amd import _native as N

    return N.default_client(0)


@pytest.fixture
def whole_array_rl(monkeypatch):
    """The plain 2-D entry point on its whole-array passes, the kernels the batched form shares (small
    2-D arrays otherwise take the one-launch tiled iteration, equal to rounding but not bit for bit)."""
    monkeypatch.setenv("LUCEM_RL_TILED", "0")
