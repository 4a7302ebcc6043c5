"""Test configuration: the package, the oracle's driver and the scenario modules on the path; the `gpu` marker for the
tests that need an MI355X (skipped, not failed, where no GPU is visible)."""

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
for p in (ROOT / "exact-ballpark_amd", ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

import torch  # noqa: E402  (before numpy users: torch's HIP runtime is the one the native library binds)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (the HIP kernels on the device)")


def pytest_collection_modifyitems(config, items):
    if torch.cuda.is_available():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)
