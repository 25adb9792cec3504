"""Shared fixtures: the HIP binding (package dir mosaic-art-rust_amd/), the
C oracle (the checker) and a session GPU context for tests marked `gpu`."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mosaic-art-rust_amd")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def _load(name, path):
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def mh():
    return _load("mosaic_hip", os.path.join(PKG, "mosaic_hip.py"))


@pytest.fixture(scope="session")
def synth():
    return _load("mosaic_synth", os.path.join(PKG, "synth.py"))


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.load()
    return oracle


@pytest.fixture(scope="session")
def ctx(mh):
    c = mh.Context(0)
    yield c
    c.close()
