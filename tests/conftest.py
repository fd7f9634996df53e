"""Shared test setup: repo root on sys.path, the `gpu` marker, helpers."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


def gpu_available() -> bool:
    """True on a machine with a visible AMD GPU (checked without initialising HIP)."""
    return os.path.exists("/dev/kfd") and any(
        n.startswith("renderD") for n in os.listdir("/dev/dri")) if os.path.isdir("/dev/dri") else False


def pytest_collection_modifyitems(config, items):
    if gpu_available():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)
