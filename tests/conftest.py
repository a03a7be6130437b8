import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def load_agx():
    """The product's ctypes binding (athena-gamma_amd/__init__.py)."""
    name = "athena_gamma_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(ROOT, "athena-gamma_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_pgen():
    """Seeded synthetic initial data of the package (athena-gamma_amd/pgen.py, numpy)."""
    name = "agx_pgen"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "athena-gamma_amd", "pgen.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_smr():
    """Host block tree and EMF-correction work lists of the package (athena-gamma_amd/smr.py)."""
    name = "agx_smr"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "athena-gamma_amd", "smr.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def agx():
    return load_agx()


@pytest.fixture(scope="session")
def orc():
    from oracle import orc as o
    o.lib()
    return o
