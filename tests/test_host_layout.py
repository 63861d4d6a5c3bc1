"""Host-side layout code (so-bogus_amd/csrc/sb_host.cpp) on the CPU: tiling
invariants, an emulation of the kernels' segmented wave scan at each tile's
recorded depth, BSR transpose, symmetric expansion, W = H Minv H^T assembly
(DualFrictionProblem::computeFrom, src/Interfaces/FrictionProblem.cpp:30-61) and
its diagonal against dense products, and both colorings (Coloring::compute,
src/Core/BlockSolvers/Coloring.impl.hpp:21-76, as written, and first fit). The checks
live in tests/cpp/test_host.cpp; integer-valued data, bit-exact comparisons."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "cpp" / "test_host.cpp"
HOST = ROOT / "so-bogus_amd" / "csrc" / "sb_host.cpp"
BIN = ROOT / "tests" / "cpp" / "_build" / "test_host"


def _binary() -> Path:
    deps = [SRC, HOST, HOST.with_suffix(".hpp")]
    if not BIN.exists() or any(d.stat().st_mtime > BIN.stat().st_mtime for d in deps):
        cxx = shutil.which("g++")
        if cxx is None:
            pytest.skip("g++ not available")
        BIN.parent.mkdir(parents=True, exist_ok=True)
        subprocess.run([cxx, "-std=c++17", "-O2", "-Wall", "-fopenmp", f"-I{HOST.parent}", str(SRC), str(HOST),
                        "-o", str(BIN)], check=True)
    return BIN


def test_host_layout():
    out = subprocess.run([str(_binary())], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.startswith("ok "), out.stdout
