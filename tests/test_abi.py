"""The C-ABI library (no GPU needed): it loads, exports every symbol that
include/sobogus_hip.h declares, and reports errors instead of crashing when no
device is present."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def L(R):
    if not R.LIB_PATH.exists():
        subprocess.run(["make", "-j8", "-C", str(ROOT / "so-bogus_amd")], check=True)
    return R.lib()


def test_header_symbols_exported(R, L):
    declared = R.header_symbols()
    assert len(declared) >= 25
    for name in declared:
        assert hasattr(L, name), name
    assert sorted(declared) == sorted(R.exported_symbols())


def test_nm_exports_c_linkage(R, L):
    out = subprocess.run(["nm", "-D", "--defined-only", str(R.LIB_PATH)], capture_output=True, text=True).stdout
    syms = set(re.findall(r"\b(sb_[a-z0-9_]+)\b", out))
    for name in R.header_symbols():
        assert name in syms, f"{name} not exported with C linkage"


def test_status_strings_and_defaults(R, L):
    assert L.sb_status_string(0) == b"ok"
    assert b"not converged" in L.sb_status_string(1)
    assert L.sb_status_string(-6).startswith(b"grid barrier")
    o = R.pg_options()
    assert (o.variant, o.law, o.tol, o.max_iters, o.ls_iters) == (R.PG_APGD, R.LAW_SOC, 1e-6, 300, 8)
    assert (o.ls_optimistic, o.ls_pessimistic, o.ls_armijo) == (1.25, 0.5, 0.5)
    g = R.gs_options()
    assert (g.law, g.strategy, g.tol, g.max_iters, g.eval_every, g.skip_iters, g.try_zero) == (
        R.LAW_COULOMB, R.REV_HYBRID, 1e-6, 250, 25, 10, 1)


def test_invalid_arguments_rejected(R, L):
    assert L.sb_ctx_create(None, 0) == -1
    assert L.sb_ctx_destroy(None) == -1
    assert L.sb_problem_destroy(None) == -1
    assert L.sb_bsr_destroy(None) == -1


def test_no_device_is_an_error_not_a_crash(R, L):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert L.sb_ctx_create(C.byref(h), 0) == -4  # SB_ERR_NO_DEVICE
