import importlib.util
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def _load_pkg():
    if "spock_amd" in sys.modules:
        return sys.modules["spock_amd"]
    pkg_dir = os.path.join(ROOT, "spock.cu_amd")
    spec = importlib.util.spec_from_file_location("spock_amd", os.path.join(pkg_dir, "__init__.py"),
                                                  submodule_search_locations=[pkg_dir])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["spock_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


_load_pkg()


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def pytest_collection_modifyitems(session, config, items):
    """GPU tests that run the C++ drop-in driver in a process of its own go first, before this
    process touches the GPU: the driver's 64-workgroup cooperative launch was refused while this
    process, still alive, had made cooperative launches of its own.  The order is otherwise kept."""
    items.sort(key=lambda it: 0 if it.get_closest_marker("gpu") and it.path.name == "test_dropin.py" else 1)


@pytest.fixture(scope="session")
def problem_dirs(tmp_path_factory):
    """Generated problem folders shared by the whole session."""
    from spock_amd.api import configs
    base = tmp_path_factory.mktemp("problems")
    out = {}
    for name in ("fixture", "small", "mid", "poly", "nocon", "affine", "pwi", "linear"):
        d = str(base / name)
        configs.generate(name, d)
        out[name] = d
    return out


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b))))
