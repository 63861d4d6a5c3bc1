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


# ---------------------------------------------------------------- matrix-free (product) GS
# ProductGaussSeidel (ProductGaussSeidel.impl.hpp:128-304) through sb_pgs_solve, on
# problems without an assembled W.

def test_small_pgs_known_solution_gpu(R, G, ctx):
    """tests/ProductGaussSeidel.cpp:44-61: from x0 = ones, res < 1e-8 and sol to 1e-4."""
    print("case: 2-contact product GS", flush=True)
    prob, sol = G.small_friction_pb()
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    g = dp.solve_pgs(np.ones(6), law=R.LAW_COULOMB, tol=1e-8)
    assert g.residual < 1e-8
    assert _is_approx(sol, g.r, 1e-4)


def test_small_pgs_trivial_diagonal_gpu(R, G, ctx):
    """tests/ProductGaussSeidel.cpp:27-39: D = identity (the operator is H H^T) with b from
    the true inverse mass; the result re-evaluated on H H^T is below 1e-8."""
    import dataclasses
    print("case: 2-contact product GS, D = I", flush=True)
    prob, _ = G.small_friction_pb()
    eye = dataclasses.replace(prob, minv=np.tile(np.eye(6).reshape(36), (prob.nb, 1)))
    dp = R.DualFrictionProblem(ctx, eye, assemble=False)
    g = dp.solve_pgs(np.ones(6), law=R.LAW_COULOMB, tol=1e-8)
    assert g.residual < 1e-8
    H = eye.dense_h()
    u = H @ H.T @ g.r + eye.b
    assert dp.eval_with(g.r, law=R.LAW_COULOMB, op=R.OP_MATRIX_FREE, scaling=R.SCALING_DIAGONAL) < 1e-8
    assert np.allclose(dp.apply_w(g.r, op=R.OP_MATRIX_FREE) + eye.b, u, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("mode", ["COLORING_REFERENCE", "COLORING_GREEDY"])
def test_coloring_without_w(R, G, ctx, mode):
    """A problem without W is colored from H's contact-body incidence, to the same coloring."""
    print(f"case: coloring from H {mode}", flush=True)
    prob = G.workload("small-1k")
    pw = R.DualFrictionProblem(ctx, prob, assemble=True).coloring(getattr(R, mode))
    ph = R.DualFrictionProblem(ctx, prob, assemble=False).coloring(getattr(R, mode))
    assert np.array_equal(pw[0], ph[0]) and np.array_equal(pw[1], ph[1])


@pytest.mark.parametrize("law", ["LAW_COULOMB", "LAW_SOC"])
def test_product_gs_parity_1k(R, O, G, ctx, law):
    print(f"case: 1k product GS {law}", flush=True)
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    perm, colors = dp.coloring(R.COLORING_GREEDY)
    dp.apply_permutation(perm)
    op = O.OracleProblem(prob, assemble=False)
    op.permute(perm.astype(np.int32))
    coul = law == "LAW_COULOMB"
    g = dp.solve_pgs(np.zeros(3 * prob.nc), law=getattr(R, law), tol=1e-8, max_iters=1000, colors=colors,
                     hist_cap=100)
    ores, ox, oinfo = op.pgs_solve(np.zeros(3 * prob.nc), tol=1e-8, max_iters=1000, coulomb=coul,
                                   ndx_ref_per_color=True, colors=colors)
    print(f"   colors {len(colors) - 1} gpu {g.iterations} it res {g.residual:.3e}; oracle {oinfo['iterations']} "
          f"res {ores:.3e}; |r-rref|/|rref| {np.linalg.norm(g.r - ox) / max(np.linalg.norm(ox), 1e-300):.2e}",
          flush=True)
    assert g.iterations == oinfo["iterations"]
    k = min(len(g.history), len(oinfo["history"]))
    assert np.allclose(g.history[:k], oinfo["history"][:k], rtol=1e-6, atol=1e-14)
    assert np.linalg.norm(g.r - ox) <= 1e-6 * max(np.linalg.norm(ox), 1e-12)
    assert g.converged and op.eval(g.r, coulomb=coul, assembled=False) <= 1.05e-8


def test_product_gs_equals_gs(R, G, ctx):
    """With the same coloring the product form and the assembled-W kernel are the same sweeps."""
    print("case: product GS vs assembled GS", flush=True)
    prob = G.generate(nc=3000, nb=1200, seed=71)
    dp = R.DualFrictionProblem(ctx, prob, assemble=True)
    perm, colors = dp.coloring(R.COLORING_GREEDY)
    dp.apply_permutation(perm)
    a = dp.solve_gs(np.zeros(3 * prob.nc), law=R.LAW_SOC, tol=1e-8, max_iters=2000, colors=colors, hist_cap=200)
    b = dp.solve_pgs(np.zeros(3 * prob.nc), law=R.LAW_SOC, tol=1e-8, max_iters=2000, colors=colors, hist_cap=200)
    print(f"   GS {a.iterations} it {a.residual:.3e}, product GS {b.iterations} it {b.residual:.3e}", flush=True)
    assert a.iterations == b.iterations
    assert np.allclose(a.history, b.history, rtol=1e-8, atol=0)
    assert np.linalg.norm(a.r - b.r) <= 1e-10 * np.linalg.norm(a.r)


def test_product_gs_50k_soc(R, O, G, ctx):
    """50,000 contacts, SOC law, tol 1e-8, no W: converged, residual re-evaluated by the
    oracle (matrix-free) below the tolerance."""
    print("case: 50k product GS SOC", flush=True)
    prob = G.workload("contacts-50k")
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    perm, colors = dp.coloring(R.COLORING_GREEDY)
    dp.apply_permutation(perm)
    g = dp.solve_pgs(np.zeros(3 * prob.nc), law=R.LAW_SOC, tol=1e-8, max_iters=3000, colors=colors)
    print(f"   {len(colors) - 1} colors, {g.iterations} sweeps, res {g.residual:.3e}, {g.device_ms:.1f} ms", flush=True)
    assert g.converged
    op = O.OracleProblem(prob, assemble=False)
    op.permute(perm.astype(np.int32))
    assert op.eval(g.r, coulomb=False, assembled=False) < 1.05e-8


def _natural_runs(prob):
    """Maximal runs of consecutive contacts that share no body (the natural sweep order)."""
    stamp = np.full(prob.nb, -1)
    runs, c = [0], 0
    for i in range(prob.nc):
        bodies = prob.colidx[prob.rowptr[i]:prob.rowptr[i + 1]]
        if np.any(stamp[bodies] == c):
            runs.append(i)
            c += 1
        stamp[bodies] = c
    runs.append(prob.nc)
    return np.array(runs)


def test_product_gs_natural_order(R, O, G, ctx):
    """Without colors the contacts are swept in their own order (the reference's
    sequential ProductGaussSeidel), in runs of contacts that share no body."""
    print("case: 1k product GS, natural order", flush=True)
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    g = dp.solve_pgs(np.zeros(3 * prob.nc), law=R.LAW_COULOMB, tol=1e-8, max_iters=1000, hist_cap=100)
    op = O.OracleProblem(prob, assemble=False)
    runs = _natural_runs(prob)
    ores, ox, oinfo = op.pgs_solve(np.zeros(3 * prob.nc), tol=1e-8, max_iters=1000, coulomb=True,
                                   ndx_ref_per_color=True, colors=runs)
    sres, sx, sinfo = op.pgs_solve(np.zeros(3 * prob.nc), tol=1e-8, max_iters=1000, coulomb=True)
    print(f"   {len(runs) - 1} runs; gpu {g.iterations} it; oracle (runs) {oinfo['iterations']}, "
          f"oracle (sequential) {sinfo['iterations']}", flush=True)
    assert g.iterations == oinfo["iterations"]
    assert np.allclose(g.history, oinfo["history"][:len(g.history)], rtol=1e-6, atol=1e-14)
    assert g.converged and sres < 1e-8
    assert np.linalg.norm(g.r - sx) <= 1e-3 * np.linalg.norm(sx)


def test_product_gs_rejects_coupled_color(R, G, ctx):
    """A color holding two contacts of one body is refused (their body updates would race)."""
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    with pytest.raises(RuntimeError):
        dp.solve_pgs(np.zeros(3 * prob.nc), law=R.LAW_COULOMB, colors=np.array([0, prob.nc]))


def test_gs_natural_order(R, O, G, ctx):
    """Assembled-W GS without colors: the contacts are swept in their own order, in
    runs of uncoupled contacts (the reference's sequential GaussSeidel)."""
    print("case: 1k GS, natural order", flush=True)
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob)
    g = dp.solve_gs(np.zeros(3 * prob.nc), law=R.LAW_COULOMB, tol=1e-8, max_iters=1000, hist_cap=100)
    op = O.OracleProblem(prob)
    runs = _natural_runs(prob)
    ores, ox, oinfo = op.gs_solve(np.zeros(3 * prob.nc), tol=1e-8, max_iters=1000, coulomb=True,
                                  ndx_ref_per_color=True, colors=runs)
    print(f"   {len(runs) - 1} runs; gpu {g.iterations} it; oracle (runs) {oinfo['iterations']}", flush=True)
    assert g.iterations == oinfo["iterations"]
    assert np.allclose(g.history, oinfo["history"][:len(g.history)], rtol=1e-6, atol=1e-14)
    assert np.linalg.norm(g.r - ox) <= 1e-6 * np.linalg.norm(ox)


def test_gs_rejects_coupled_color(R, G, ctx):
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob)
    with pytest.raises(RuntimeError):
        dp.solve_gs(np.zeros(3 * prob.nc), law=R.LAW_COULOMB, colors=np.array([0, prob.nc]))


# ---------------------------------------------------------------- mixed-precision GS
# FP32 off-diagonal gathers (W32, x32) with the FP64 local solve and FB evaluation, then
# FP64 sweeps from the FP32 stage's best iterate. Bar: converged, the returned r
# re-evaluated by the FP64 oracle meets the tolerance (x1.05), and r agrees with the
# FP64 GS solution to the tolerance scale (the SOC problem's solution is unique).

@pytest.mark.parametrize("law", ["LAW_COULOMB", "LAW_SOC"])
def test_mixed_gs_1k(R, O, G, ctx, law):
    print(f"case: 1k mixed-precision GS {law}", flush=True)
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob)
    perm, colors = dp.coloring(R.COLORING_GREEDY)
    dp.apply_permutation(perm)
    op = O.OracleProblem(prob)
    op.permute(perm.astype(np.int32))
    coul = law == "LAW_COULOMB"
    kw = dict(law=getattr(R, law), tol=1e-8, max_iters=2000, colors=colors)
    m = dp.solve_gs(np.zeros(3 * prob.nc), precision=R.PRECISION_MIXED, **kw)
    d = dp.solve_gs(np.zeros(3 * prob.nc), precision=R.PRECISION_FP64, **kw)
    print(f"   mixed {m.iterations} sweeps ({m.iterations - m.refine_iterations} FP32, FP32 res "
          f"{m.mixed_residual:.2e}), res {m.residual:.3e}; FP64 {d.iterations} sweeps", flush=True)
    assert m.converged and d.converged
    assert m.iterations - m.refine_iterations > 0 and m.refine_iterations >= 1
    assert op.eval(m.r, coulomb=coul) <= 1.05e-8
    assert m.residual == pytest.approx(op.eval(m.r, coulomb=coul), rel=1e-6)
    if not coul:
        assert np.linalg.norm(m.r - d.r) <= 1e-3 * np.linalg.norm(d.r)


def test_mixed_gs_50k_soc(R, O, G, ctx):
    """50,000 contacts, SOC law, tol 1e-8, assembled W, 16 greedy colors: the mixed GS
    converges to an oracle-verified result; its time is printed beside the FP64 GS."""
    print("case: 50k mixed-precision GS SOC", flush=True)
    prob = G.workload("contacts-50k")
    dp = R.DualFrictionProblem(ctx, prob)
    perm, colors = dp.coloring(R.COLORING_GREEDY)
    dp.apply_permutation(perm)
    kw = dict(law=R.LAW_SOC, tol=1e-8, max_iters=3000, colors=colors)
    m = dp.solve_gs(np.zeros(3 * prob.nc), precision=R.PRECISION_MIXED, **kw)
    d = dp.solve_gs(np.zeros(3 * prob.nc), precision=R.PRECISION_FP64, **kw)
    print(f"   mixed {m.iterations} sweeps ({m.iterations - m.refine_iterations} FP32, FP32 res {m.mixed_residual:.2e}) "
          f"{m.device_ms:.1f} ms; FP64 {d.iterations} sweeps {d.device_ms:.1f} ms", flush=True)
    assert m.converged and d.converged
    op = O.OracleProblem(prob)
    op.permute(perm.astype(np.int32))
    assert op.eval(m.r, coulomb=False) <= 1.05e-8
    assert np.linalg.norm(m.r - d.r) <= 1e-3 * np.linalg.norm(d.r)


def test_mixed_gs_refused_where_unsupported(R, G, ctx):
    prob = G.workload("small-1k")
    dp = R.DualFrictionProblem(ctx, prob, assemble=False)
    with pytest.raises(R.SbError) as e:
        dp.solve_pgs(np.zeros(3 * prob.nc), law=R.LAW_SOC, tol=1e-8, precision=R.PRECISION_MIXED)
    assert e.value.code == R.SB_ERR_UNSUPPORTED
