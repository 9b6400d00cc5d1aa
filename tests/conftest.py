"""Shared fixtures.  GPU tests carry @pytest.mark.gpu (registered in
pytest.ini) and run only with ``-m gpu`` on a machine with an MI355X; they
fail, rather than skip, when no HIP device is present."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) HIP device")


def load_golden(name: str) -> np.ndarray:
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as o
    o.build()
    return o


@pytest.fixture(scope="session")
def hip():
    """The HIP library on a live gfx950 device (GPU tests only)."""
    import torch
    from scir_amd import _lib
    L = _lib.lib()
    assert torch.cuda.is_available(), "GPU test run without a visible HIP device"
    torch.cuda.init()
    _lib.check(L.scir_hip_init())
    return _lib


def rng(seed: int) -> np.random.Generator:
    return np.random.default_rng(seed)
