"""Shared pytest configuration.

Makes the package (``high-performance-lbm-on-gpu-_amd/lbm``) and the oracle importable,
registers the ``gpu`` and ``slow`` markers, and skips ``gpu`` tests when no ROCm
device is visible, so the same suite runs on CPU-only machines.
"""

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
PACKAGE_DIR = ROOT / "high-performance-lbm-on-gpu-_amd"
for path in (str(PACKAGE_DIR), str(ROOT)):
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and liblbm_hip.so (skipped otherwise)")
    config.addinivalue_line("markers", "slow: long validation runs")


def pytest_collection_modifyitems(config, items):
    from lbm.backends import gpu_available

    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no ROCm GPU visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
