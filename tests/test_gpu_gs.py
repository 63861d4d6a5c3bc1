"""Colored Gauss-Seidel on the GPU against the oracle.

The reference's own assertions (tests/GaussSeidel.cpp:19-65) pin the 2-contact
fixture. On generated problems the GPU and the oracle run the same coloring and
the per-color step-size reference (ndxRef), so the sweeps are the same algorithm:
residual histories agree to rounding and the results agree tightly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _is_approx(a, b, prec):
    return np.linalg.norm(a - b) <= prec * min(np.linalg.norm(a), np.linalg.norm(b))


@pytest.mark.parametrize("strategy", ["HYBRID", "PURE_NEWTON", "PURE_ENUMERATIVE", "REV_HYBRID"])
def test_small_gs_coulomb_gpu(R, G, ctx, strategy):
    print(f"case: 2-contact GS {strategy}", flush=True)
    prob, sol = G.small_friction_pb()
    dp = R.DualFrictionProblem(ctx, prob)
    g = dp.solve_gs(np.ones(6), law=R.LAW_COULOMB, strategy=getattr(R, strategy))
    assert g.residual < 1e-8
    assert _is_approx(sol, g.r, 1e-4)


def test_small_gs_soc_gpu(R, G, ctx):
    print("case: 2-contact GS SOC", flush=True)
    prob, sol = G.small_friction_pb()
    dp = R.DualFrictionProblem(ctx, prob)
    g = dp.solve_gs(np.ones(6), law=R.LAW_SOC, tol=1e-8)
    assert g.residual < 1e-8


@pytest.mark.parametrize("mode", ["COLORING_GREEDY", "COLORING_REFERENCE"])
def test_coloring_matches_oracle(R, O, G, ctx, mode):
    print(f"case: coloring {mode}", flush=True)
    prob = G.generate(nc=2000, nb=1000, seed=41)
    dp = R.DualFrictionProblem(ctx, prob)
    perm, colors = dp.coloring(getattr(R, mode))
    assert np.array_equal(np.sort(perm), np.arange(prob.nc))
    rp, ci, _ = dp.w_lower()
    color_of = np.empty(prob.nc, int)
    for c in range(len(colors) - 1):
        color_of[perm[colors[c]:colors[c + 1]]] = c
    for i in range(prob.nc):
        for p in range(rp[i], rp[i + 1]):
            if ci[p] != i:
                assert color_of[i] != color_of[ci[p]]
    if mode == "COLORING_REFERENCE":
        op = O.OracleProblem(prob)
        operm, ocolors = op.coloring()
        assert np.array_equal(perm, operm) and np.array_equal(colors, ocolors)


@pytest.mark.parametrize("law", ["LAW_COULOMB", "LAW_SOC"])
def test_colored_gs_parity_1k(R, O, G, ctx, law):
    print(f"case: 1k colored GS {law}", flush=True)
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob)
    perm, colors = dp.coloring(R.COLORING_GREEDY)
    dp.apply_permutation(perm)
    op = O.OracleProblem(prob)
    op.permute(perm.astype(np.int32))
    coul = law == "LAW_COULOMB"
    g = dp.solve_gs(np.zeros(3 * prob.nc), law=getattr(R, law), tol=1e-8, max_iters=1000, colors=colors,
                    hist_cap=100)
    ores, ox, oinfo = op.gs_solve(np.zeros(3 * prob.nc), tol=1e-8, max_iters=1000, coulomb=coul,
                                  ndx_ref_per_color=True, colors=colors)
    print(f"   colors {len(colors) - 1} gpu {g.iterations} it res {g.residual:.3e}; oracle {oinfo['iterations']} "
          f"res {ores:.3e}; |r-rref|/|rref| {np.linalg.norm(g.r - ox) / max(np.linalg.norm(ox), 1e-300):.2e}",
          flush=True)
    assert g.iterations == oinfo["iterations"]
    k = min(len(g.history), len(oinfo["history"]))
    assert np.allclose(g.history[:k], oinfo["history"][:k], rtol=1e-6, atol=1e-14)
    assert np.linalg.norm(g.r - ox) <= 1e-6 * max(np.linalg.norm(ox), 1e-12)
    if g.converged:
        assert op.eval(g.r, coulomb=coul) <= 1.05e-8


def test_gs_iteration_cap_returns_best(R, O, G, ctx):
    print("case: GS max_iters / x_best", flush=True)
    prob = G.generate(nc=600, nb=250, seed=43)
    dp = R.DualFrictionProblem(ctx, prob)
    perm, colors = dp.coloring(R.COLORING_GREEDY)
    dp.apply_permutation(perm)
    g = dp.solve_gs(np.zeros(3 * prob.nc), law=R.LAW_COULOMB, tol=1e-30, max_iters=50, colors=colors)
    assert g.status == R.SB_NOT_CONVERGED
    assert dp.eval_with(g.r, law=R.LAW_COULOMB) == pytest.approx(g.residual, rel=1e-9)
