"""Shared test setup: import paths, the `gpu` marker and device fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "khrr-nbody_amd")
ORACLE = os.path.join(ROOT, "oracle")
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (PKG, ORACLE, ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C-ABI on cuda:0)")


@pytest.fixture(scope="session")
def ctx():
    """One library context on cuda:0 for the GPU tests (null stream, which
    is also torch's default stream, so torch copies and khrr kernels are
    ordered)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import khrr
    c = khrr.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")
