import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
PKG_DIR = os.path.join(ROOT, "potentialframework.jl_amd")


def load_package():
    """Import potentialframework.jl_amd (a directory name with a dot) as ``pfw_amd``."""
    if "pfw_amd" in sys.modules:
        return sys.modules["pfw_amd"]
    spec = importlib.util.spec_from_file_location("pfw_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pfw_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (HIP device)")


@pytest.fixture(scope="session")
def pfw():
    return load_package()
