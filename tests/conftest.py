"""Test configuration: the `gpu` marker and in-tree imports.

CPU tests (`-m "not gpu"`) pin the oracle and check the C-ABI library loads;
GPU tests (`-m gpu`) run the HIP path through the C-ABI and compare it with
the oracle on the same seeded inputs.
"""
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def _ensure_built():
    from code_amd import build
    if not (build.BUILD / build.LIB_NAME).exists():
        build.build_hip()
    if not (build.ORACLE_BUILD / build.ORACLE_LIB).exists():
        build.build_oracle()


_ensure_built()


@pytest.fixture(scope="session")
def device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch.device("cuda:0")
