"""Host-side layout code (so-bogus_amd/csrc/sb_host.cpp) on the CPU: tiling
invariants, an emulation of the kernels' segmented wave scan at each tile's
recorded depth, BSR transpose, symmetric expansion, W = H Minv H^T assembly
(DualFrictionProblem::computeFrom, src/Interfaces/FrictionProblem.cpp:30-61) and
its diagonal against dense products, and both colorings (Coloring::compute,
src/Core/BlockSolvers/Coloring.impl.hpp:21-76, as written, and first fit). The checks
live in tests/cpp/test_host.cpp; integer-valued data, bit-exact comparisons.
The same checks also run in a build under AddressSanitizer and
UndefinedBehaviorSanitizer (host code only, g++), every finding fatal."""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "cpp" / "test_host.cpp"
HOST = ROOT / "so-bogus_amd" / "csrc" / "sb_host.cpp"
BIN = ROOT / "tests" / "cpp" / "_build" / "test_host"
BIN_SAN = ROOT / "tests" / "cpp" / "_build" / "test_host_asan_ubsan"
SAN_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]


def _binary(out: Path, flags: list[str]) -> Path:
    deps = [SRC, HOST, HOST.with_suffix(".hpp")]
    if not out.exists() or any(d.stat().st_mtime > out.stat().st_mtime for d in deps):
        cxx = shutil.which("g++")
        if cxx is None:
            pytest.skip("g++ not available")
        out.parent.mkdir(parents=True, exist_ok=True)
        subprocess.run([cxx, "-std=c++17", *flags, "-Wall", "-fopenmp", f"-I{HOST.parent}", str(SRC), str(HOST),
                        "-o", str(out)], check=True)
    return out


def test_host_layout():
    out = subprocess.run([str(_binary(BIN, ["-O2"]))], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.startswith("ok "), out.stdout


def test_host_layout_sanitized():
    """Every host layout check under ASan + UBSan: out-of-bounds accesses, leaks,
    signed overflow and misaligned or null accesses abort the run."""
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    out = subprocess.run([str(_binary(BIN_SAN, SAN_FLAGS))], capture_output=True, text=True, timeout=300, env=env)
    assert out.returncode == 0, out.stderr[-4000:]
    assert out.stdout.startswith("ok "), out.stdout
    assert "runtime error" not in out.stderr, out.stderr[-4000:]
