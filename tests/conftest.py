import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_pkg():
    if "mdoodz_amd" in sys.modules:
        return sys.modules["mdoodz_amd"]
    spec = importlib.util.spec_from_file_location("mdoodz_amd", os.path.join(ROOT, "mdoodz6.0_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules["mdoodz_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def mdz():
    return load_pkg()


@pytest.fixture(scope="session")
def orc(mdz):
    import oracle_bind
    return oracle_bind.load(mdz)


@pytest.fixture(scope="session")
def hot(mdz):
    return mdz.lib_hot()
