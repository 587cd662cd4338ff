import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_ibm():
    """The product's Python binding (ib.cu_amd/ibm.py)."""
    return _load("ibm", os.path.join(ROOT, "ib.cu_amd", "ibm.py"))


def load_inputs():
    return _load("ibm_inputs", os.path.join(ROOT, "ib.cu_amd", "inputs.py"))


def load_surfaces():
    """Host-side RBF operator construction (ib.cu_amd/surfaces.py)."""
    return _load("ibm_surfaces", os.path.join(ROOT, "ib.cu_amd", "surfaces.py"))


def load_orc_rbf():
    return _load("orc_rbf", os.path.join(ROOT, "oracle", "rbf.py"))


def load_orc():
    """The CPU oracle binding (oracle/orc.py)."""
    return _load("orc", os.path.join(ROOT, "oracle", "orc.py"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def ibm():
    return load_ibm()


@pytest.fixture(scope="session")
def inputs():
    return load_inputs()


@pytest.fixture(scope="session")
def orc():
    return load_orc()


@pytest.fixture(scope="session")
def ctx(ibm):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    c = ibm.Context(0)
    yield c
    c.close()
