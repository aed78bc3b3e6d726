"""Shared pytest setup: the `gpu` marker and import paths."""

import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "adaptive-regression-trees-art-_amd"
for path in (str(ROOT), str(PKG), str(Path(__file__).resolve().parent)):
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libart_hip.so")


@pytest.fixture(scope="session")
def torch_gpu():
    """torch with a visible HIP device and the built library, or fail loudly."""
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device visible")
    from art import _hip

    _hip.load()  # raises when libart_hip.so is missing: no silent fallback
    return torch
