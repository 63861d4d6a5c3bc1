"""Projected-gradient solves on the GPU against the oracle.

Parity bar (FP64 path): the GPU result's FB residual, recomputed by the oracle,
is <= tol (x1.05 for the different summation order of the residual itself), and
the result agrees with the oracle running the same algorithm from the same
start: relative |r - r_ref| <= 1e-6 for small problems where the trajectories
stay together. Iteration counts are compared loosely (line-search decisions can
flip on rounding, after which the iterates legitimately differ).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small(G):
    return G.small_friction_pb()


@pytest.mark.parametrize("variant", ["PG_APGD", "PG_DESCENT", "PG_STANDARD", "PG_CONJUGATED", "PG_SPG"])
def test_small_pg_soc_gpu(R, ctx, small, variant):
    """tests/ProjectedGradient.cpp:15-60 on the GPU path, every variant: residual < 1e-8."""
    print(f"case: 2-contact PG {variant}", flush=True)
    prob, sol = small
    dp = R.DualFrictionProblem(ctx, prob)
    res = dp.solve_pg(np.ones(6), variant=getattr(R, variant), law=R.LAW_SOC, tol=1e-8)
    assert res.residual < 1e-8 and res.converged


def test_small_pg_matrix_free_gpu(R, ctx, small):
    """tests/ProjectedGradient.cpp:63-83: matrix-free H Minv H^T product, residual < 1e-8."""
    print("case: 2-contact matrix-free APGD", flush=True)
    prob, sol = small
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    res = dp.solve_pg(np.ones(6), law=R.LAW_SOC, op=R.OP_MATRIX_FREE, tol=1e-8)
    assert res.residual < 1e-8


@pytest.mark.parametrize("variant", ["PG_APGD", "PG_CONJUGATED"])
def test_small_pg_trajectory_matches_oracle(R, O, ctx, small, variant):
    """Same algorithm, same start: residual history and result match the oracle.
    (Standard is not iterate-exact anywhere: the first trial of its line search is the
    Cauchy step along the projected gradient, where the Armijo test holds with equality
    in exact arithmetic whenever the projection is inactive, so rounding decides it;
    the oracle and a numpy restatement of the reference already part at iteration 1.)"""
    print(f"case: 2-contact {variant} trajectory", flush=True)
    prob, sol = small
    dp = R.DualFrictionProblem(ctx, prob)
    op = O.OracleProblem(prob)
    g = dp.solve_pg(np.ones(6), variant=getattr(R, variant), law=R.LAW_SOC, tol=1e-8, hist_cap=1000)
    ores, ox, oinfo = op.pg_solve(np.ones(6), variant=getattr(O, variant), tol=1e-8, coulomb=False)
    assert g.iterations == oinfo["iterations"]
    assert np.allclose(g.history, oinfo["history"], rtol=1e-6, atol=1e-15)
    assert np.allclose(g.r, ox, rtol=1e-8, atol=1e-12)


@pytest.mark.parametrize("variant", ["PG_APGD", "PG_SPG", "PG_STANDARD", "PG_CONJUGATED"])
@pytest.mark.parametrize("law", ["LAW_COULOMB", "LAW_SOC"])
def test_1k_parity(R, O, G, ctx, variant, law):
    """1,000 contacts / 500 bodies, tol 1e-8, cold start."""
    print(f"case: 1k {variant} {law}", flush=True)
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob)
    op = O.OracleProblem(prob)
    coul = law == "LAW_COULOMB"
    g = dp.solve_pg(np.zeros(3 * prob.nc), variant=getattr(R, variant), law=getattr(R, law), tol=1e-8,
                    max_iters=5000, hist_cap=5000)
    assert g.converged
    chk = op.eval(g.r, coulomb=coul)
    assert chk <= 1.05e-8
    ores, ox, oinfo = op.pg_solve(np.zeros(3 * prob.nc), variant=getattr(O, variant), tol=1e-8, max_iters=5000,
                                  coulomb=coul)
    # same algorithm: iteration counts within 25%, and the first iterations coincide
    # (Standard: only the starting residual - its line search's first trial is an exact
    # tie of the Armijo test, decided by rounding)
    assert abs(g.iterations - oinfo["iterations"]) <= 0.25 * oinfo["iterations"] + 5
    k = min(1 if variant == "PG_STANDARD" else 20, len(g.history), len(oinfo["history"]))
    assert np.allclose(g.history[:k], oinfo["history"][:k], rtol=1e-6)
    # the velocities u = W r + b agree to the tolerance scale
    u_g = op.apply_w(g.r) + prob.b
    u_o = op.apply_w(ox) + prob.b
    assert np.linalg.norm(u_g - u_o) / np.linalg.norm(u_o) < 1e-2
    print(f"   gpu iters {g.iterations} oracle {oinfo['iterations']} res {g.residual:.3e} check {chk:.3e} "
          f"|r-rref|/|rref| {np.linalg.norm(g.r - ox) / np.linalg.norm(ox):.2e}", flush=True)


def test_matrix_free_equals_assembled_operator(R, O, G, ctx):
    """APGD with the matrix-free product and diagonal scaling reaches the same
    tolerance as with the assembled W (both checked by the oracle)."""
    print("case: 1k APGD matrix-free vs assembled", flush=True)
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob)
    op = O.OracleProblem(prob)
    g = dp.solve_pg(np.zeros(3 * prob.nc), law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL,
                    tol=1e-8, max_iters=5000)
    assert g.converged and op.eval(g.r, coulomb=True) <= 1.05e-8


@pytest.mark.parametrize("minv_case", ["spd", "static_body", "asymmetric"])
def test_matrix_free_factored_operator(R, O, G, ctx, minv_case):
    """The matrix-free solvers apply W = F F^T with F = H L (Minv = L L^T, factored on the
    device); a problem with a body whose Minv is not positive definite (a static body,
    Minv = 0) or not symmetric keeps the unfactored product H (Minv (H^T x)). Either way
    the GPU follows the oracle's matrix-free APGD (full Minv blocks) over the first
    iterations and meets the tolerance."""
    print(f"case: 600-contact matrix-free APGD, Minv {minv_case}", flush=True)
    prob = G.generate(nc=600, nb=300, seed=31)
    if minv_case == "static_body":
        prob.minv[0] = 0.0
    elif minv_case == "asymmetric":
        m = prob.minv[5].reshape(6, 6)
        m[0, 1] += 0.05 * m[0, 0]  # still positive definite; the lower triangle alone is not Minv
        prob.minv[5] = m.reshape(36)
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    op = O.OracleProblem(prob)
    g = dp.solve_pg(np.zeros(3 * prob.nc), law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, tol=1e-8, max_iters=5000,
                    hist_cap=5000)
    assert g.converged
    assert op.eval(g.r, coulomb=True, diag_scaling=False, assembled=False) <= 1.05e-8
    ores, ox, oinfo = op.pg_solve(np.zeros(3 * prob.nc), tol=1e-8, max_iters=5000, coulomb=True, assembled=False)
    k = min(20, len(g.history), len(oinfo["history"]))
    assert np.allclose(g.history[:k], oinfo["history"][:k], rtol=1e-6)


def test_warm_start_and_iteration_cap(R, O, G, ctx):
    print("case: warm start / max_iters", flush=True)
    prob = G.generate(nc=500, nb=250, seed=17)
    dp = R.DualFrictionProblem(ctx, prob)
    g = dp.solve_pg(np.zeros(3 * prob.nc), law=R.LAW_COULOMB, tol=1e-8, max_iters=5000)
    assert g.converged
    again = dp.solve_pg(g.r, law=R.LAW_COULOMB, tol=1e-8, max_iters=5000)
    assert again.converged and again.iterations == 0
    capped = dp.solve_pg(np.zeros(3 * prob.nc), law=R.LAW_COULOMB, tol=1e-14, max_iters=7)
    assert capped.status == R.SB_NOT_CONVERGED and capped.iterations == 7
    # the returned r is the best iterate: its residual is the returned value
    assert dp.eval_with(capped.r, law=R.LAW_COULOMB) == pytest.approx(capped.residual, rel=1e-9)


def test_device_resident_solve(R, O, G, ctx):
    print("case: device-resident r", flush=True)
    prob = G.generate(nc=800, nb=400, seed=23)
    dp = R.DualFrictionProblem(ctx, prob)
    r = R.DeviceArray(ctx, 3 * prob.nc).upload(np.zeros(3 * prob.nc))
    g = dp.solve_pg(r, law=R.LAW_COULOMB, tol=1e-8, max_iters=5000, on_device=True)
    host = dp.solve_pg(np.zeros(3 * prob.nc), law=R.LAW_COULOMB, tol=1e-8, max_iters=5000)
    assert g.converged
    assert np.array_equal(r.download(), host.r)  # deterministic: same bits both ways


@pytest.mark.parametrize("variant", ["PG_STANDARD", "PG_CONJUGATED"])
def test_standard_conjugated_matrix_free_and_mixed(R, O, G, ctx, variant):
    """Standard and Conjugated on the matrix-free operator (FP64 and mixed precision):
    converged, every result re-evaluated by the oracle below 1.05 tol."""
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    op = O.OracleProblem(prob, assemble=False)
    for prec in ("PRECISION_FP64", "PRECISION_MIXED"):
        print(f"case: 1k {variant} matrix-free {prec}", flush=True)
        g = dp.solve_pg(np.zeros(3 * prob.nc), variant=getattr(R, variant), law=R.LAW_COULOMB,
                        op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL, tol=1e-8, max_iters=5000,
                        precision=getattr(R, prec))
        assert g.converged
        assert op.eval(g.r, coulomb=True, assembled=False, diag_scaling=True) <= 1.05e-8


def test_batch_refuses_standard_and_conjugated(R, G, ctx):
    """The batch kernel runs APGD / Descent / SPG only; the other variants are refused."""
    B = R.Batch(ctx, G.batch(3, lo=64, hi=200))
    for v in (R.PG_STANDARD, R.PG_CONJUGATED):
        with pytest.raises(R.SbError) as e:
            B.solve(variant=v, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, tol=1e-6)
        assert e.value.code == -5


@pytest.mark.parametrize("law", ["LAW_COULOMB", "LAW_SOC"])
def test_standard_parity_without_the_armijo_tie(R, O, G, ctx, law):
    """Standard with an Armijo coefficient of 0.4 (set in both codes): its first line-search
    trial is then no longer an exact tie, and the GPU follows the oracle's residual history
    and iteration count, as APGD does with the defaults."""
    print(f"case: 1k PG_STANDARD armijo 0.4 {law}", flush=True)
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob)
    op = O.OracleProblem(prob)
    coul = law == "LAW_COULOMB"
    g = dp.solve_pg(np.zeros(3 * prob.nc), variant=R.PG_STANDARD, law=getattr(R, law), tol=1e-8, max_iters=5000,
                    ls_armijo=0.4, hist_cap=5000)
    O.set_pg_line_search(armijo=0.4)
    try:
        ores, ox, oinfo = op.pg_solve(np.zeros(3 * prob.nc), variant=O.PG_STANDARD, tol=1e-8, max_iters=5000,
                                      coulomb=coul)
    finally:
        O.set_pg_line_search()
    print(f"   gpu iters {g.iterations} oracle {oinfo['iterations']}", flush=True)
    assert g.converged and op.eval(g.r, coulomb=coul) <= 1.05e-8
    assert abs(g.iterations - oinfo["iterations"]) <= 0.25 * oinfo["iterations"] + 5
    k = min(20, len(g.history), len(oinfo["history"]))
    assert np.allclose(g.history[:k], oinfo["history"][:k], rtol=1e-6)
