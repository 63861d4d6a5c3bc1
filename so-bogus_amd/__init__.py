"""so-bogus_amd — MI355X (gfx950) Coulomb-friction solve path of So-bogus.

The directory name is not a Python identifier; load it with ``load_package()``
from the repository root (it registers the package as ``sobogus_amd``).
"""
from __future__ import annotations

import importlib.util
import sys
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
__all__ = ["generator", "runtime"]


def load_package():
    """Import this directory as the package ``sobogus_amd`` and return it."""
    if "sobogus_amd" in sys.modules:
        return sys.modules["sobogus_amd"]
    spec = importlib.util.spec_from_file_location("sobogus_amd", PKG_DIR / "__init__.py",
                                                  submodule_search_locations=[str(PKG_DIR)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["sobogus_amd"] = mod
    spec.loader.exec_module(mod)
    return mod
