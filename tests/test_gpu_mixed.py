"""Mixed-precision projected gradient on the GPU: FP32 matrices, iterates, products
and projection with FP64 residual, dot products and decisions, then an FP64 stage
warm-started from the FP32 result that verifies the tolerance (and refines when
it is not met).

Parity bar: the returned r, re-evaluated by the FP64 oracle, meets the tolerance
(x1.05 for the residual's own summation order); the velocities u = W r + b agree
with the FP64 GPU solve's to the tolerance scale (rel. 1e-2: iterates of a
tolerance-level solve are not unique to better than that); the FP32 stage's own
residual is reported and within 1e-3 relative of the FP64 re-evaluation of the
same iterate, or the FP64 stage iterated.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _u(prob, r):
    W = prob.dense_w()
    return W @ r + prob.b


@pytest.mark.parametrize("variant", ["PG_APGD", "PG_SPG"])
@pytest.mark.parametrize("op", ["OP_ASSEMBLED", "OP_MATRIX_FREE"])
def test_mixed_matches_fp64_1k(G, R, O, ctx, variant, op):
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob, assemble=True)
    op_ = O.OracleProblem(prob)
    kw = dict(variant=getattr(R, variant), law=R.LAW_COULOMB, op=getattr(R, op), scaling=R.SCALING_DIAGONAL,
              tol=1e-8, max_iters=5000)
    print(f"[mixed] 1k {variant} {op}", flush=True)
    m = dp.solve_pg(np.zeros(3 * prob.nc), precision=R.PRECISION_MIXED, **kw)
    d = dp.solve_pg(np.zeros(3 * prob.nc), precision=R.PRECISION_FP64, **kw)
    assert m.converged and d.converged
    chk = op_.eval(m.r, coulomb=True)
    assert chk <= 1.05e-8
    assert m.residual == pytest.approx(chk, rel=1e-6)
    # the FP32 stage reached tol * 0.95 (the default factor) on its own residual, or the FP64 stage refined
    assert m.mixed_residual < 0.95e-8 or m.refine_iterations > 0
    assert m.refine_spmvs >= 1  # the FP64 verification is at least one operator application
    u_m, u_d = _u(prob, m.r), _u(prob, d.r)
    assert np.linalg.norm(u_m - u_d) / np.linalg.norm(u_d) < 1e-2


def test_mixed_small_reference_problem(G, R, ctx):
    prob, sol = G.small_friction_pb()
    dp = R.DualFrictionProblem(ctx, prob, assemble=True)
    for variant in (R.PG_APGD, R.PG_DESCENT, R.PG_SPG):
        res = dp.solve_pg(np.ones(6), variant=variant, law=R.LAW_SOC, tol=1e-8, precision=R.PRECISION_MIXED)
        assert res.residual < 1e-8 and res.converged


def test_mixed_50k_matrix_free(G, R, O, ctx):
    prob = G.workload("contacts-50k")
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    print("[mixed] 50k matrix-free APGD tol 1e-6", flush=True)
    m = dp.solve_pg(np.zeros(3 * prob.nc), variant=R.PG_APGD, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE,
                    scaling=R.SCALING_DIAGONAL, tol=1e-6, max_iters=20000, precision=R.PRECISION_MIXED)
    assert m.converged
    chk = O.OracleProblem(prob).eval(m.r, coulomb=True)  # assembled W, diagonal scaling
    assert chk <= 1.05e-6


def test_mixed_warm_start_verifies_without_iterating(G, R, ctx):
    """A converged FP64 result passed back in: the FP32 stage stops at once or
    iterates, and the FP64 stage verifies with one operator application."""
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob, assemble=True)
    kw = dict(law=R.LAW_COULOMB, tol=1e-8, max_iters=5000)
    d = dp.solve_pg(np.zeros(3 * prob.nc), **kw)
    again = dp.solve_pg(d.r, precision=R.PRECISION_FP64, **kw)
    assert again.converged and again.iterations == 0 and again.spmvs == 1
