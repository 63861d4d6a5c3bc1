"""sb_pg_solve_noalloc through the C ABI against sb_pg_solve.

The allocation-free variant keeps the persistent launches' grid state in the
library's fixed per-GPU array instead of the context's buffers; the arithmetic is
the same, so its results must equal sb_pg_solve's bit for bit (FP64 and mixed
precision, assembled and matrix-free), on two calls in a row with the second on a
smaller problem, and after a device-side value update (the factored operator
tiles refreshed in place).
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [("OP_ASSEMBLED", "PRECISION_FP64"), ("OP_ASSEMBLED", "PRECISION_MIXED"),
         ("OP_MATRIX_FREE", "PRECISION_FP64"), ("OP_MATRIX_FREE", "PRECISION_MIXED")]


def _opts(R, op, prec, tol=1e-8):
    return R.pg_options(variant=R.PG_APGD, law=R.LAW_COULOMB, op=getattr(R, op), scaling=R.SCALING_DIAGONAL,
                        tol=tol, max_iters=3000, precision=getattr(R, prec))


def _pair(R, ctx, prob, opts):
    """(sb_pg_solve result, sb_pg_solve_noalloc result), each on a device r from r0 = 0."""
    dp = R.DualFrictionProblem(ctx, prob, assemble=True)
    ra = R.DeviceArray(ctx, 3 * prob.nc).zero()
    a = dp.solve_pg(ra, opts, on_device=True)
    rb = R.DeviceArray(ctx, 3 * prob.nc).zero()
    b = dp.solve_pg_noalloc(rb, opts)
    return a, ra.download(), b, rb.download()


@pytest.mark.parametrize("op,prec", CASES)
def test_noalloc_equals_pg_solve_two_calls(R, G, ctx, op, prec):
    print(f"[noalloc] {op} {prec}: 1000 then 400 contacts", flush=True)
    opts = _opts(R, op, prec)
    for prob in (G.workload("small-1k"), G.generate(nc=400, nb=200, seed=11)):
        a, xa, b, xb = _pair(R, ctx, prob, opts)
        assert a.converged and b.converged
        assert b.status == a.status and b.iterations == a.iterations and b.spmvs == a.spmvs
        assert b.residual == a.residual
        assert np.array_equal(xa, xb)


def test_noalloc_after_device_update(R, G, ctx):
    """After sb_problem_update_values the matrix-free solve recomputes its factored tiles
    in place; the result equals sb_pg_solve's on an identically updated problem."""
    print("[noalloc] matrix-free after a device value update", flush=True)
    prob = G.workload("small-1k")
    opts = _opts(R, "OP_MATRIX_FREE", "PRECISION_FP64")
    newp = dataclasses.replace(prob, hvals=prob.hvals * 1.05, minv=prob.minv * 0.9, b=prob.b * 0.8)
    out = []
    for noalloc in (False, True):
        dp = R.DualFrictionProblem(ctx, prob, assemble=False)
        dp.solve_pg(np.zeros(3 * prob.nc), opts)  # allocates the factored tiles
        dp.update_values(newp.hvals, newp.minv, newp.b, None)
        r = R.DeviceArray(ctx, 3 * prob.nc).zero()
        res = dp.solve_pg_noalloc(r, opts) if noalloc else dp.solve_pg(r, opts, on_device=True)
        out.append((res, r.download()))
    (a, xa), (b, xb) = out
    assert a.converged and b.iterations == a.iterations and b.residual == a.residual
    assert np.array_equal(xa, xb)
    fresh = R.DualFrictionProblem(ctx, newp, assemble=False)
    assert fresh.eval_with(xb, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL) < 1.05e-8


def test_noalloc_refusals(R, G, ctx):
    """A matrix-free solve before the factored tiles exist and a request for a residual
    history are refused (nothing is allocated to serve them)."""
    print("[noalloc] refusals", flush=True)
    prob = G.generate(nc=200, nb=100, seed=13)
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    r = R.DeviceArray(ctx, 3 * prob.nc).zero()
    with pytest.raises(R.SbError) as e:
        dp.solve_pg_noalloc(r, _opts(R, "OP_MATRIX_FREE", "PRECISION_FP64"))
    assert e.value.code == R.SB_ERR_LIMIT
    o = _opts(R, "OP_MATRIX_FREE", "PRECISION_FP64")
    dp.solve_pg(np.zeros(3 * prob.nc), o)
    st = R.Stats()
    hist = np.zeros(16)
    st.history = hist.ctypes.data_as(C.POINTER(C.c_double))
    st.history_capacity = 16
    res = C.c_double()
    code = R.lib().sb_pg_solve_noalloc(ctx.h, dp.h, C.byref(o), r.ptr, C.byref(res), C.byref(st))
    assert code == R.SB_ERR_INVALID
    ok = dp.solve_pg_noalloc(r, o)
    assert ok.converged
