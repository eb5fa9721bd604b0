"""Shared test setup: import paths and the `gpu` marker.

GPU tests are marked `@pytest.mark.gpu`; the CPU suite runs with -m "not gpu".
"""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "pfc-models_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")
    config.addinivalue_line("markers", "slow: longer CPU oracle runs")


@pytest.fixture(scope="session")
def hip_device():
    """Skip unless a HIP device is visible; on a GPU box the library must load."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from pfc_hip import _lib
    _lib.load()  # raises if the built library is missing: no silent fallback
    return torch.device("cuda", 0)
