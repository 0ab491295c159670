import pathlib
import sys

import pytest

REPO = pathlib.Path(__file__).resolve().parent.parent
PKG = REPO / "factor-cuda_amd"
for p in (str(REPO), str(PKG), str(REPO / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libfc_hip.so")


def pytest_collection_modifyitems(config, items):
    # gpu tests first within a file keep the default order; nothing to reorder
    return


@pytest.fixture(scope="session")
def rng():
    import numpy as np
    return np.random.default_rng(12345)


@pytest.fixture(autouse=True, scope="session")
def _torch_hip_first(request):
    """GPU tests call libfc_hip through ctypes and torch in one process: torch
    creates its HIP context before the library's first call, whichever test
    file runs first (the other order leaves torch without a device)."""
    if any(it.get_closest_marker("gpu") for it in request.session.items):
        try:
            import torch
            if torch.cuda.is_available():
                torch.cuda.init()
        except Exception:  # a missing GPU is reported by the tests themselves
            pass
    yield
