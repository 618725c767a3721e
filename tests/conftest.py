import importlib.util
import os
import sys

import numpy as np
import pytest
# torch first: libpearl_gpu.so then binds to the HIP runtime torch loads, so
# tests that use torch on the GPU and the library share one runtime
import torch  # noqa: F401

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG_DIR = os.path.join(ROOT, "fe-c_eutectoid_phase-field_simulation_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU reference comparison")


def load_pkg():
    """Import the solver package (its directory name is not an identifier)."""
    if "pearl_amd" in sys.modules:
        return sys.modules["pearl_amd"]
    spec = importlib.util.spec_from_file_location("pearl_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["pearl_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def load_golden(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    keys = [str(k) for k in d["param_keys"]]
    d["pd"] = dict(zip(keys, d["params"].tolist()))
    return d


@pytest.fixture(scope="session")
def pkg():
    return load_pkg()


@pytest.fixture(scope="session")
def O():
    import oracle
    oracle.build()
    return oracle
