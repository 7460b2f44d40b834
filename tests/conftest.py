import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_ibim():
    """Import the package directory ibim-levelset_amd/ as module ibim_amd."""
    if "ibim_amd" in sys.modules:
        return sys.modules["ibim_amd"]
    path = os.path.join(ROOT, "ibim-levelset_amd", "__init__.py")
    spec = importlib.util.spec_from_file_location("ibim_amd", path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules["ibim_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def ibim():
    return load_ibim()


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    return torch.device("cuda:0")
