"""Shared fixtures.  Tests marked ``gpu`` need a ROCm device and the built
HIP library; everything else runs on the CPU (oracle, host logic, ABI)."""

from __future__ import annotations

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def device():
    """The GPU under test; the HIP library must load (no fallback)."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU visible")
    from warp_amd import _native

    _native.lib()  # raises if the library is missing
    return torch.device("cuda", 0)


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle

    oracle.lib()
    return oracle
