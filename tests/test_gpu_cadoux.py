"""Cadoux fixed point on the GPU (sb_pg_solve_cadoux / sb_gs_solve_cadoux) against the
oracle's restatement of src/Interfaces/Cadoux.hpp:44-155.

With the colored GS inner solver on the same coloring the inner solves are the
oracle's sweep for sweep, so the fixed-point histories agree to rounding. With APGD
inner solves (whose iteration counts depend on rounding) the results agree at the
tolerance: every returned r is re-evaluated by the oracle on the Coulomb law.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _is_approx(a, b, prec):
    return np.linalg.norm(a - b) <= prec * min(np.linalg.norm(a), np.linalg.norm(b))


@pytest.mark.parametrize("solver", ["gs", "pg"])
def test_small_cadoux_gpu(R, G, O, ctx, solver):
    print(f"case: 2-contact Cadoux {solver}", flush=True)
    prob, sol = G.small_friction_pb()
    dp = R.DualFrictionProblem(ctx, prob)
    kw = dict(tol=1e-8) if solver == "gs" else dict(tol=1e-8, max_iters=1000)
    c = dp.solve_cadoux(np.zeros(6), solver=solver, iterations=50, hist_cap=60, **kw)
    print(f"   {c.iterations} fixed-point iterations, residual {c.residual:.3e}", flush=True)
    assert c.converged and c.residual < 1e-8
    assert _is_approx(sol, c.r, 1e-4)
    assert len(c.history) == c.iterations + 1
    op = O.OracleProblem(prob)
    assert op.eval(c.r, coulomb=True) == pytest.approx(c.residual, rel=1e-9)


@pytest.mark.parametrize("solver", ["gs", "pg"])
def test_small_cadoux_velocity_gpu(R, G, O, ctx, solver):
    print(f"case: 2-contact CadouxVel {solver}", flush=True)
    prob, _ = G.small_friction_pb()
    dp = R.DualFrictionProblem(ctx, prob)
    kw = dict(tol=1e-8) if solver == "gs" else dict(tol=1e-8, max_iters=1000)
    c = dp.solve_cadoux(np.zeros(6), solver=solver, iterations=100, velocity=True, hist_cap=110, **kw)
    op = O.OracleProblem(prob)
    ores, ou, od = op.cadoux_solve(np.zeros(6), use_gs=solver == "gs", tol=1e-8, iterations=100, velocity=True,
                                   inner_max_iters=kw.get("max_iters", 250))
    print(f"   gpu {c.iterations} it {c.residual:.3e}; oracle {od['iterations']} it {ores:.3e}", flush=True)
    assert c.converged and c.residual < 1e-8
    assert np.linalg.norm(c.r - ou) <= 1e-4 * np.linalg.norm(ou)


def test_cadoux_gs_parity_1k(R, O, G, ctx):
    print("case: 1k Cadoux, colored GS inner", flush=True)
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob)
    perm, colors = dp.coloring(R.COLORING_GREEDY)
    dp.apply_permutation(perm)
    op = O.OracleProblem(prob)
    op.permute(perm.astype(np.int32))
    c = dp.solve_cadoux(np.zeros(3 * prob.nc), solver="gs", colors=colors, iterations=100, tol=1e-6,
                        max_iters=2000, hist_cap=120)
    ores, ox, od = op.cadoux_solve(np.zeros(3 * prob.nc), use_gs=True, tol=1e-6, iterations=100,
                                   inner_max_iters=2000, colors=colors)
    print(f"   gpu {c.iterations} it {c.residual:.3e} ({c.spmvs} applications); oracle {od['iterations']} it "
          f"{ores:.3e}", flush=True)
    assert c.converged and c.iterations == od["iterations"]
    assert np.allclose(c.history, od["history"], rtol=1e-6, atol=1e-14)
    assert np.linalg.norm(c.r - ox) <= 1e-6 * np.linalg.norm(ox)
    assert op.eval(c.r, coulomb=True) < 1.05e-6


@pytest.mark.parametrize("precision", ["PRECISION_FP64", "PRECISION_MIXED"])
@pytest.mark.parametrize("op_name", ["OP_ASSEMBLED", "OP_MATRIX_FREE"])
def test_cadoux_pg_1k(R, O, G, ctx, precision, op_name):
    print(f"case: 1k Cadoux, APGD inner {precision} {op_name}", flush=True)
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob)
    c = dp.solve_cadoux(np.zeros(3 * prob.nc), solver="pg", iterations=100, tol=1e-6, max_iters=3000,
                        op=getattr(R, op_name), scaling=R.SCALING_DIAGONAL, precision=getattr(R, precision),
                        hist_cap=120)
    op = O.OracleProblem(prob)
    ores, ox, od = op.cadoux_solve(np.zeros(3 * prob.nc), use_gs=False, tol=1e-6, iterations=100,
                                   inner_max_iters=3000, assembled=op_name == "OP_ASSEMBLED")
    chk = op.eval(c.r, coulomb=True)
    print(f"   gpu {c.iterations} it {c.residual:.3e} (oracle-evaluated {chk:.3e}); oracle {od['iterations']} it "
          f"{ores:.3e}", flush=True)
    assert c.converged and chk < 1.05e-6
    assert chk == pytest.approx(c.residual, rel=1e-6)
    assert abs(c.iterations - od["iterations"]) <= max(3, od["iterations"] // 2)


def test_cadoux_restores_b(R, G, ctx):
    """The problem's affine term is the caller's again after a fixed-point solve."""
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob)
    r = np.random.default_rng(3).standard_normal(3 * prob.nc) * 1e-2
    before = dp.eval_with(r, law=R.LAW_COULOMB)
    dp.solve_cadoux(np.zeros(3 * prob.nc), solver="pg", iterations=3, tol=1e-6)
    assert dp.eval_with(r, law=R.LAW_COULOMB) == before


def test_cadoux_gs_after_device_update_keeps_b(R, G, ctx):
    """After a device-side value update the host copies lag the device; a GS Cadoux
    solve (whose set-up reads the host copies) must not take an inner problem's
    affine term for the problem's b: a later host-side update of mu alone uploads
    the caller's b, and residuals match a freshly created problem's."""
    print("case: Cadoux GS after sb_problem_update_values, then sb_problem_update(mu)", flush=True)
    import dataclasses
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob)
    newb = prob.b * 0.75 + 0.1
    dp.update_values(b=newb)
    dp.solve_cadoux(np.zeros(3 * prob.nc), solver="gs", iterations=3, tol=1e-6, max_iters=50)
    newmu = prob.mu * 0.9
    dp.update(mu=newmu)
    fresh = R.DualFrictionProblem(ctx, dataclasses.replace(prob, b=newb, mu=newmu))
    r = np.random.default_rng(4).standard_normal(3 * prob.nc) * 1e-2
    for law in (R.LAW_COULOMB, R.LAW_SOC):
        assert dp.eval_with(r, law=law) == fresh.eval_with(r, law=law)
    # u = W r + b from the residual at r = 0 on the SOC law sees b alone
    assert dp.eval_with(np.zeros(3 * prob.nc), law=R.LAW_SOC) == fresh.eval_with(np.zeros(3 * prob.nc),
                                                                                  law=R.LAW_SOC)
