"""Shared pytest configuration.

Markers: `gpu` = needs a ROCm GPU and the built HIP library (libvgi.so).
GPU tests fail loudly (not skip) on a machine with a GPU whose library is missing, so a
silent fallback is impossible.
"""
from __future__ import annotations

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: test needs a ROCm GPU (MI355X) and libvgi.so")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def _gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _gpu_available():
        return
    skip = pytest.mark.skip(reason="no ROCm GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def oracle64():
    import oracle

    return oracle.Oracle(64)


@pytest.fixture(scope="session")
def oracle32():
    import oracle

    return oracle.Oracle(32)


@pytest.fixture(scope="session")
def vgi():
    """The HIP library, loaded (raises if not built)."""
    from violet_amd import _lib, gpu

    gpu.require_gpu()
    return _lib.lib()
