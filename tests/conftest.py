import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "torch: uses PyTorch tensors on the GPU (the autograd harness)")


@pytest.hookimpl(trylast=True)  # after -m/-k deselection
def pytest_collection_modifyitems(config, items):
    # The torch harness tests share the process with the other GPU tests, and
    # torch and libalp_hip.so must share one HIP runtime: when any of them is
    # selected, load torch (its runtime) before libalp_hip.so is first loaded.
    if any(item.get_closest_marker("gpu") and "torch" in item.keywords for item in items):
        import torch  # noqa: F401


@pytest.fixture(scope="session")
def golden():
    import numpy as np

    def load(name):
        return np.load(os.path.join(GOLDEN, name + ".npz"))

    return load
