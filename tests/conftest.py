"""Shared fixtures: the package (loaded from so-bogus_amd/), the oracle, a GPU context."""
from __future__ import annotations

import importlib.util
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "so-bogus_amd"
sys.path.insert(0, str(ROOT / "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")


def load_package():
    if "sobogus_amd" in sys.modules:
        return sys.modules["sobogus_amd"]
    spec = importlib.util.spec_from_file_location("sobogus_amd", PKG / "__init__.py",
                                                  submodule_search_locations=[str(PKG)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["sobogus_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


load_package()


@pytest.fixture(scope="session")
def G():
    from sobogus_amd import generator
    return generator


@pytest.fixture(scope="session")
def R():
    from sobogus_amd import runtime
    return runtime


@pytest.fixture(scope="session")
def O():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def ctx(R):
    c = R.Context(0)
    print(f"\n[gpu] {c.arch} with {c.n_cu} CUs", flush=True)
    yield c
    c.close()
