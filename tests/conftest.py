import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

_RUNS_GPU_TESTS = False


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built librinnd_hip.so")


def pytest_collection_modifyitems(session, config, items):
    global _RUNS_GPU_TESTS
    _RUNS_GPU_TESTS = any(item.get_closest_marker("gpu") for item in items)


@pytest.fixture(scope="session", autouse=True)
def _torch_runtime_first():
    """PyTorch ships its own HIP runtime; it finds no device if it initialises after librinnd_hip.so's runtime has,
    so a session with GPU tests brings torch's device up first (tests use torch only for device buffers and the
    FP64 recall).  Autouse session fixtures run before the session's other fixtures (the `ctx` below)."""
    if not _RUNS_GPU_TESTS:
        return
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


@pytest.fixture(scope="session")
def ctx():
    from rinnd_amd import _lib
    return _lib.Context.default(0)
