"""Shared test setup: import paths, the `gpu` marker, seeded RNGs.

GPU tests (marked `gpu`) call the HIP operators through tcuscan_ops / the C
ABI and compare with the CPU oracle in oracle/oracle.py.  They fail loudly
when no GPU is visible rather than skipping, so a GPU run can never pass on
a silent fallback.
"""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "pytorch-tcuscan_amd")
for p in (REPO, PKG, os.path.join(REPO, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture
def dev():
    import torch

    assert torch.cuda.is_available(), "GPU test run without a visible GPU"
    return torch.device("cuda:0")


@pytest.fixture
def ops():
    import torch  # noqa: F401

    import tcuscan_ops

    return tcuscan_ops


@pytest.fixture
def rng():
    return np.random.default_rng(42)
