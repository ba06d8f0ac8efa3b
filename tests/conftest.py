import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def _load_package():
    if "flame3d_amd" in sys.modules:
        return sys.modules["flame3d_amd"]
    pkg_dir = os.path.join(ROOT, "flame3d.jl_amd")
    spec = importlib.util.spec_from_file_location(
        "flame3d_amd", os.path.join(pkg_dir, "__init__.py"),
        submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["flame3d_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


_load_package()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def pkg():
    """The flame3d_amd package with its abi and surrogate modules loaded."""
    import flame3d_amd.abi  # noqa: F401
    import flame3d_amd.surrogate  # noqa: F401
    return sys.modules["flame3d_amd"]
