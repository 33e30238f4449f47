import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); parity tests")


def _has_gpu() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _has_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU in this environment")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(scope="session")
def dev():
    import torch
    return torch.device("cuda:0")


@pytest.fixture(autouse=True)
def _device_status_clean(request):
    """Every GPU test ends with no device status flag raised (a radix
    look-back that gave up, inexact int64 coordinates), unless it raises
    them on purpose and clears them itself."""
    yield
    if "gpu" not in request.keywords or not _has_gpu():
        return
    from parlllelrtree_amd import rtree
    flags = rtree.device_status(clear=True)
    assert flags == 0, f"device status flags 0x{flags:x} raised during {request.node.name}"


def pytest_report_header(config):
    return f"prt tests, GPU available: {_has_gpu()}, cwd {os.getcwd()}"
