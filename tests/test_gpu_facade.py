"""The reference's solver tests written against the C++ facade (tests/cpp/test_facade.cpp),
run on the GPU path."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
BIN = ROOT / "tests" / "cpp" / "_build" / "test_facade"

pytestmark = pytest.mark.gpu


def test_cpp_facade_reference_tests():
    print("case: C++ facade (BlockMV, GaussSeidel, ProjectedGradient)", flush=True)
    assert BIN.exists(), "build the facade tests first (make -C so-bogus_amd)"
    p = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
