"""Test configuration: the `gpu` marker, and the built libraries.

CPU tests (`-m "not gpu"`) exercise the oracles and the host logic; GPU tests
(`-m gpu`) run the HIP engine on an MI355X and compare it with the oracles.
"""

import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X and the built HIP engine")


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle

    oracle.build()
    return oracle.library()


@pytest.fixture(scope="session")
def engine():
    """The built HIP engine on a GPU; a GPU test with no engine fails loudly."""
    from glazier_amd import engine as eng

    eng.library()
    return eng
