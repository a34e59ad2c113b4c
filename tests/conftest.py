"""pytest setup: the `gpu` marker, the package loader, and the oracle path.

The product package lives in `tensorcrossinterpolation.jl_amd/` (not a Python
identifier) and is loaded under the name `tci_amd`.  The oracle (CPU
restatement, test infrastructure) lives in `oracle/`.
"""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "tensorcrossinterpolation.jl_amd")
for p in (os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "engine_host"),
          os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def load_pkg():
    if "tci_amd" in sys.modules:
        return sys.modules["tci_amd"]
    spec = importlib.util.spec_from_file_location("tci_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["tci_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) and the built libtci_hip.so")


@pytest.fixture(scope="session")
def tci():
    return load_pkg()


@pytest.fixture(scope="session")
def ctx(tci):
    # torch is only the plumbing for device buffers in a few tests; let it
    # find the device before the library's own HIP calls
    import torch
    if torch.cuda.is_available():
        torch.cuda.init()
    return tci.default_context()
